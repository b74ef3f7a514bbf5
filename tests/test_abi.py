"""CPU-side checks of the C ABI: the library builds, loads and exports every
entry point include/g5.h declares; calls that need a GPU fail cleanly."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(name="g5.h"):
    txt = open(os.path.join(REPO, "include", name)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int|void)\s+(g5_\w+)\s*\(", txt, re.M)))


def all_header_functions():
    """Every C entry point declared in any header under include/."""
    names = set()
    for f in os.listdir(os.path.join(REPO, "include")):
        if f.endswith(".h"):
            names |= set(header_functions(f))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    import graph500_amd
    from graph500_amd import build
    if not os.path.exists(build.LIB):
        build.build()
    return graph500_amd.load_library()


def test_header_lists_every_binding():
    from graph500_amd import g5
    assert header_functions() == sorted(g5.EXPORTS)


def test_library_exports_every_header_symbol(lib):
    from graph500_amd import g5
    assert all_header_functions() == sorted(g5.EXPORTS + g5.MEASURE_EXPORTS)
    for name in all_header_functions():
        assert hasattr(lib, name), name


def test_measure_gather_rejects_bad_sizes(lib):
    ms = ctypes.c_double(0)
    chk = ctypes.c_uint64(0)
    for nbytes, gathers in ((12, 1), (0, 1), (64, 0)):
        assert lib.g5_measure_gather(0, nbytes, gathers, ctypes.byref(ms), ctypes.byref(chk)) == -1


def test_version_and_device_query_without_gpu(lib):
    assert b"gfx950" in lib.g5_version()
    n = ctypes.c_int(-1)
    assert lib.g5_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0


def test_null_context_is_rejected(lib):
    assert lib.g5_destroy(None) == 0
    assert lib.g5_run_bfs(None, 0) == -1
    assert lib.g5_construct(None, 2, 0) == -1
    assert lib.g5_last_error(None) == b"null context"


def test_create_fails_cleanly_on_missing_device(lib):
    h = ctypes.c_void_p()
    assert lib.g5_create(10_000, ctypes.byref(h)) == -1
    assert not h.value


def test_context_requires_gpu():
    import graph500_amd
    n = ctypes.c_int(0)
    graph500_amd.load_library().g5_device_count(ctypes.byref(n))
    if n.value == 0:
        with pytest.raises(RuntimeError):
            graph500_amd.G5Context(0)


def test_cpp_shim_compiles_against_reference_shaped_edge_list(tmp_path):
    """include/g5_bfs.hpp compiles stand-alone with an edge list that has the
    EdgeListStorage read interface (generator/graph_generator.hpp:98-180)."""
    import subprocess
    src = tmp_path / "shim.cc"
    src.write_text(r'''
#include "g5_bfs.hpp"
struct Tuple { uint32_t a, b, c; };
struct List {
  typedef Tuple edge_type;
  int beginRead(bool) { return 1; }
  int read(Tuple** p) { static Tuple t[2]; *p = t; return 0; }
  void endRead() {}
};
void f(g5::BfsOnGPU& b, List* l, int64_t* pred) {
  b.construct(l);
  b.prepare_bfs();
  int64_t roots[4];
  b.find_roots(roots, 4);
  if (b.graph_.has_edge(roots[0])) b.run_bfs(roots[0], pred);
  b.get_pred(pred);
  int64_t n = 0;
  (void)b.validate(roots[0], pred, &n);
  (void)b.graph_.pred_size();
  b.end_bfs();
}
''')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(REPO, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_driver_reports_missing_gpu():
    """g5bench fails with a message (exit 1), not a crash, without a device."""
    import subprocess
    from graph500_amd import build
    if not os.path.exists(build.DRIVER):
        build.build()
    n = ctypes.c_int(0)
    import graph500_amd
    graph500_amd.load_library().g5_device_count(ctypes.byref(n))
    if n.value:
        pytest.skip("a GPU is present")
    r = subprocess.run([build.DRIVER, "10"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "g5_create" in r.stderr


STRUCTS = {"GraphInfo": "g5_graph_info", "BfsStats": "g5_bfs_stats", "BatchStats": "g5_batch_stats",
           "BatchTuning": "g5_batch_tuning", "ServerConfig": "g5_server_config", "Query": "g5_query",
           "QueryResult": "g5_query_result", "ServerReport": "g5_server_report",
           "ServerGraphStats": "g5_server_graph_stats"}


def test_binding_struct_layouts_match_header(tmp_path):
    """Every ctypes structure has the field offsets and size of its C struct."""
    import subprocess
    from graph500_amd import g5
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "g5.h"', "int main(void) {"]
    expect = []
    for py, c in STRUCTS.items():
        S = getattr(g5, py)
        for name, _ in S._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (c, name, c, name))
            expect.append("%s.%s %d" % (c, name, getattr(S, name).offset))
        lines.append('  printf("%s.size %%zu\\n", sizeof(%s));' % (c, c))
        expect.append("%s.size %d" % (c, ctypes.sizeof(S)))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")
    assert [l for l in got if l] == expect
