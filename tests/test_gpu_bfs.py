"""GPU parity of the HIP path (through the C ABI) against the oracle and the
reference golden fixtures.

Bar: bit-exact for tuples, CSR, roots, levels, reached sets, direction
schedules, frontier sizes and traversed-edge counts; parents may differ from
the reference's but must pass the validator rules (mpi/validate.hpp:506-802).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return json.load(open(os.path.join(GOLD, name)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import graph500_amd
    c = graph500_amd.G5Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def s16(ctx, oracle_mod):
    say("generate+construct SCALE 16")
    ctx.generate_kronecker(16, 16, 2, 3)
    edges = ctx.edges_download()
    ctx.construct(2)
    ctx.prepare_bfs()
    return edges, oracle_mod.Csr(oracle_mod.generate_edges(16))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_generator_bit_exact_s16(s16, oracle_mod):
    edges, _ = s16
    g = gold("ref_s16_16roots.json")
    t = edges.view(np.uint32).reshape(-1, 3)
    assert sha(t) == g["tuples_sha256"]
    assert np.array_equal(edges, oracle_mod.generate_edges(16))


@pytest.mark.parametrize("scale,seeds,init", [(12, (7, 11), (5700, 1900)),
                                              (10, (2**40 + 5, 3), (5500, 100)),
                                              (1, (2, 3), (5700, 1900))])
def test_generator_other_seeds_and_initiators(oracle_mod, scale, seeds, init):
    import graph500_amd
    say("generator", scale, seeds, init)
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(scale, 16, seeds[0], seeds[1], init[0], init[1])
        got = c.edges_download()
    exp = oracle_mod.generate_edges(scale, 16, seeds[0], seeds[1], init[0], init[1])
    assert np.array_equal(got, exp)


def test_csr_byte_equal_s16(ctx, s16):
    _, oc = s16
    info = ctx.graph_info()
    g = gold("ref_s16_16roots.json")
    assert info["scale_estimate"] == 16
    assert info["num_active"] == g["num_active"] == oc.num_active
    assert info["nnz"] == g["nnz"] == oc.nnz
    assert info["directed_before"] == oc.directed_before
    assert info["num_self_loops"] == oc.num_self_loops
    a = oc.arrays()
    d = ctx.graph_download()
    assert np.array_equal(d["new2old"], a["new2old"])
    assert np.array_equal(d["offsets"], a["offsets"])
    assert np.array_equal(d["cols"], a["cols"])
    na = oc.num_active
    w = a["degree_raw"][:na] + 2 * a["self_loops"][:na]
    assert np.array_equal(d["weight"], w)
    assert info["max_used_vertex"] == g["max_used_vertex"] == oc.max_used_vertex()


def test_find_roots_match_reference(ctx, s16):
    assert ctx.find_roots(16).tolist() == gold("ref_s16_16roots.json")["roots"]
    assert ctx.find_roots(64).tolist() == gold("ref_s16_64roots.json")["roots"]


def test_has_edge_matches_oracle(ctx, s16):
    _, oc = s16
    for v in [0, 1, 5, 100, 29348, 65535, 65536, -1]:
        assert ctx.has_edge(v) == (0 <= v < oc.num_verts and oc.has_edge(v))


def test_bfs_64_roots_bit_exact_s16(ctx, s16, oracle_mod):
    edges, oc = s16
    g = gold("ref_s16_64roots.json")
    ng = oc.max_used_vertex() + 1
    for pr in g["per_root"]:
        say("bfs s16 root", pr["root"])
        ctx.run_bfs(pr["root"])
        lv = ctx.get_levels()
        assert sha(lv) == pr["levels_sha256"], pr["root"]
        st = ctx.bfs_stats()
        assert "".join("T" if d else "B" for d in st["direction"]) == pr["directions"]
        assert st["nq"] == pr["nq"]
        assert ctx.teps_edges() == pr["nedge"] == 1048570
        pred = ctx.get_pred()
        reached = pred != -1
        assert np.array_equal(reached, lv != 0xFF)
        assert np.array_equal((pred[reached] >> 48), lv[reached].astype(np.int64))
        ok, cnt, msg = oracle_mod.validate(edges, ng, pr["root"], pred)
        assert ok, msg
        assert cnt == pr["nedge"]


@pytest.mark.parametrize("scale,tuning", [(18, {}), (20, {}), (22, {}), (24, {}), (20, {"bin_edges": 1}),
                                          (22, {"bin_edges": 1 << 22}), (20, {"td_pieces": 0}),
                                          (20, {"td_pieces": 1}), (22, {"td_pieces": 7}),
                                          (22, {"td_pieces": 4096}), (20, {"block_shift": 0}),
                                          (20, {"block_shift": 17}), (22, {"block_shift": 12}),
                                          (20, {"td_head": 1}), (22, {"td_head": 16}), (22, {"td_head": 0}),
                                          (22, {"td_head": 1 << 30}), (22, {"td_head": 4, "block_shift": 14}),
                                          (22, {"td_head": 4, "td_pieces": 64}),
                                          (20, {"td_head": 4, "seg_max": 8}),
                                          (22, {"td_head": 4, "td_pieces": 64, "seg_max": 1000}),
                                          (22, {"td_head": 0, "block_shift": 14, "seg_max": 2000}),
                                          (20, {"small_edges": 0}), (22, {"small_edges": 1 << 20}),
                                          (20, {"small_edges": 64}), (24, {"small_edges": 0}),
                                          (22, {"tail_pull": 0}), (24, {"tail_pull": 0, "small_edges": 0}),
                                          (20, {"tail_pull": 64}), (24, {"tail_pull": 1})])
def test_bfs_reference_fixtures_larger_scales(scale, tuning, oracle_mod):
    """An empty tuning keeps the defaults; bin_edges 1 bins every bulk
    top-down level; td_pieces 0 takes bulk rows in row order, 1 through a
    one-piece-per-row segment list, 4096 is the largest piece count (levels
    whose segment list would not fit fall back to row order); block_shift 0
    turns target blocks off, 12 at SCALE 22 indexes only the first rows, so
    both indexed and whole-row segments occur.  seg_max lowers the segment
    count a segment-list level may hold, so that hub levels whose frontier
    rows (suffix mode) or pieces exceed it take the row-order fallback.
    small_edges 0 runs every level through the per-level path, 2^20 fuses
    every top-down level with at most 4096 frontier vertices, 64 makes the
    fused launch stop at most frontiers (too many edges), often before its
    first level.  tail_pull 0 expands every top-down level after the
    bottom-up phase top-down; the default 4 finds with the bottom-up kernel
    those whose frontier holds more than a quarter as many vertices as
    remain unvisited, 1 only those whose frontier outnumbers them, 64
    nearly all."""
    import graph500_amd
    g = gold("ref_s%d_16roots.json" % scale)
    say("scale", scale, "tuning", tuning)
    with graph500_amd.G5Context(0) as c:
        if tuning:
            c.bfs_set_tuning(**tuning)
            got = c.bfs_tuning()
            assert all(got[k] == v for k, v in tuning.items())
        c.generate_kronecker(scale, 16, 2, 3)
        t = c.edges_download().view(np.uint32).reshape(-1, 3)
        assert sha(t) == g["tuples_sha256"]
        c.construct(2)
        info = c.graph_info()
        assert info["num_active"] == g["num_active"]
        assert info["nnz"] == g["nnz"]
        assert c.find_roots(16).tolist() == g["roots"]
        c.prepare_bfs()
        for i, pr in enumerate(g["per_root"]):
            c.run_bfs(pr["root"])
            lv = c.get_levels()
            assert sha(lv) == pr["levels_sha256"], (scale, pr["root"])
            st = c.bfs_stats()
            assert "".join("T" if d else "B" for d in st["direction"]) == pr["directions"]
            assert st["nq"] == pr["nq"]
            # the bottom-up kernel runs every bottom-up level, and top-down
            # ones only after the bottom-up phase when tail_pull hands them over
            kern = st["bu_kernel"]
            first_bu = st["direction"].index(0) if 0 in st["direction"] else len(kern)
            assert all(k == 1 for k, d in zip(kern, st["direction"]) if d == 0)
            assert not any(kern[:first_bu])
            if tuning.get("tail_pull") == 0:
                assert kern == [1 - d for d in st["direction"]]
            assert c.teps_edges() == pr["nedge"]
            ok, errs = c.validate()  # every root's parents, on the GPU
            assert ok, (pr["root"], errs)
            if i < 2 and scale < 24:  # the CPU validator too, where it takes seconds
                edges = c.edges_download()
                ok, cnt, msg = oracle_mod.validate(edges, info["max_used_vertex"] + 1, pr["root"],
                                                   c.get_pred())
                assert ok, msg
                assert cnt == pr["nedge"]


def test_bfs_reference_64_roots_s24():
    """SCALE 24 with the 64 roots of the reference's REAL_BENCHMARK run
    (tests/golden/ref_s24_64roots.json): roots, level hashes, schedules,
    frontier sizes and traversed-edge counts of every root; parents
    validated on the GPU."""
    import graph500_amd
    g = gold("ref_s24_64roots.json")
    say("scale 24, 64 reference roots")
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(24, 16, 2, 3)
        c.construct(2, release_edges=True)
        info = c.graph_info()
        assert info["num_active"] == g["num_active"] and info["nnz"] == g["nnz"]
        assert c.find_roots(64).tolist() == g["roots"]
        c.prepare_bfs()
        for pr in g["per_root"]:
            c.run_bfs(pr["root"])
            assert sha(c.get_levels()) == pr["levels_sha256"], pr["root"]
            st = c.bfs_stats()
            assert "".join("T" if d else "B" for d in st["direction"]) == pr["directions"], pr["root"]
            assert st["nq"] == pr["nq"]
            assert c.teps_edges() == pr["nedge"] == 268432547
            ok, errs = c.validate()
            assert ok, (pr["root"], errs)


def _run_small(oracle_mod, v0, v1, roots, modes=(2, 1, 0)):
    import graph500_amd
    e = oracle_mod.pack_edges(v0, v1)
    oc = oracle_mod.Csr(e)
    for mode in modes:
        with graph500_amd.G5Context(0) as c:
            c.edges_append(e[: len(e) // 2])
            c.edges_append(e[len(e) // 2:])
            c.construct(mode)
            c.prepare_bfs()
            for r in roots:
                if oc.has_edge(int(r)):
                    lv_o, _, tr = oc.bfs(int(r))
                else:  # a root without edges is the whole tree
                    lv_o = np.full(oc.num_verts, 0xFF, np.uint8)
                    lv_o[int(r)] = 0
                c.run_bfs(int(r))
                lv = c.get_levels()
                assert np.array_equal(lv, lv_o), (mode, r)
                pred = c.get_pred()
                if oc.has_edge(int(r)):
                    # the validator only accepts roots up to the largest vertex with an edge
                    ok, cnt, msg = oracle_mod.validate(e, oc.max_used_vertex() + 1, int(r), pred)
                    assert ok, (mode, r, msg)
                    st = c.bfs_stats()
                    assert st["nq"] == tr["nq"], (mode, r)
                    assert c.teps_edges() == cnt == tr["teps_edges"], (mode, r)
                else:
                    assert pred[int(r)] == int(r) and (pred != -1).sum() == 1


def test_small_graphs_edge_cases(oracle_mod):
    rng = np.random.default_rng(5)
    # two components, self loops and duplicates
    v0 = np.array([0, 1, 2, 2, 5, 5, 6, 7, 7, 9, 3, 3])
    v1 = np.array([1, 2, 0, 2, 6, 6, 7, 5, 7, 9, 3, 4])
    say("small: components")
    _run_small(oracle_mod, v0, v1, [0, 5, 3, 9, 8])
    # star with a long tail (hub row split across the grid)
    n = 5000
    v0 = np.concatenate([np.zeros(n - 1, np.int64), np.arange(n, n + 60)])
    v1 = np.concatenate([np.arange(1, n), np.arange(n + 1, n + 61)])
    say("small: star+path")
    _run_small(oracle_mod, v0, v1, [0, 17, n, n + 30])
    # random sparse graph
    v0 = rng.integers(0, 3000, 20000)
    v1 = rng.integers(0, 3000, 20000)
    say("small: random")
    _run_small(oracle_mod, v0, v1, [int(v0[0]), int(v1[5])])


def test_path_graph_depth(oracle_mod):
    n = 250
    say("path graph")
    _run_small(oracle_mod, np.arange(n - 1), np.arange(1, n), [0, 125], modes=(2,))


def test_too_deep_graph_raises(oracle_mod):
    import graph500_amd
    n = 400
    e = oracle_mod.pack_edges(np.arange(n - 1), np.arange(1, n))
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(2)
        c.prepare_bfs()
        with pytest.raises(graph500_amd.G5Error):
            c.run_bfs(0)


def test_error_paths():
    import graph500_amd
    E = graph500_amd.G5Error
    with graph500_amd.G5Context(0) as c:
        with pytest.raises(E):
            c.prepare_bfs()  # no graph
        with pytest.raises(E):
            c.construct(2)  # empty edge list
        with pytest.raises(E):
            c.generate_kronecker(0)  # scale out of range
        with pytest.raises(E):
            c.bfs_set_tuning(-1)  # negative threshold
        with pytest.raises(E):
            c.bfs_set_tuning(td_pieces=4097)  # more pieces than allowed
        with pytest.raises(E):
            c.bfs_set_tuning(block_shift=5)  # blocks below 2^12 ids
        with pytest.raises(E):
            c.bfs_set_tuning(seg_max=1 << 24)  # beyond the segment count field
        with pytest.raises(E):
            c.bfs_set_tuning(small_edges=(1 << 20) + 1)  # beyond the largest fused level
        with pytest.raises(E):
            c.bfs_set_tuning(small_edges=-1)
        with pytest.raises(E):
            c.bfs_set_tuning(tail_pull=65)
        c.generate_kronecker(8)
        with pytest.raises(E):
            c.construct(7)  # unknown reorder mode
        c.construct(2)
        with pytest.raises(E):
            c.run_bfs(0)  # not prepared
        c.prepare_bfs()
        with pytest.raises(E):
            c.get_pred()  # no result yet
        with pytest.raises(E):
            c.run_bfs(1 << 8)  # root out of range
        with pytest.raises(E):
            c.run_bfs(-1)
        r = int(c.find_roots(1)[0])
        c.run_bfs(r)
        assert c.get_levels()[r] == 0


def test_model_bytes_match_oracle(ctx, s16, oracle_mod):
    """The GPU evaluation of the byte model equals the oracle's count."""
    _, oc = s16
    for r in oc.find_roots(16):
        _, _, tr = oc.bfs(int(r), want_pred=False)
        ctx.run_bfs(int(r))
        per, tot = ctx.model_bytes()
        assert per == tr["bytes"], int(r)
        assert tot == tr["total_bytes"]


def test_model_bytes_match_oracle_s20(oracle_mod):
    import graph500_amd
    say("model bytes s20")
    oc = oracle_mod.Csr(oracle_mod.generate_edges(20))
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(20)
        c.construct(2, release_edges=True)
        c.prepare_bfs()
        for r in oc.find_roots(4):
            _, _, tr = oc.bfs(int(r), want_pred=False)
            c.run_bfs(int(r))
            per, tot = c.model_bytes()
            assert per == tr["bytes"] and tot == tr["total_bytes"]


@pytest.mark.parametrize("extra", [[], ["--host-edges"]])
def test_driver_g5bench_s16(extra):
    """The C++ driver over the BfsOnGPU shim runs the main.cc flow: 16 roots,
    every pred array valid, traversed edges = pf_nedge[16] (mpi/utils.hpp:3664)."""
    import subprocess
    from graph500_amd import build
    say("g5bench 16", extra)
    r = subprocess.run([build.DRIVER, "16", "--roots", "16", *extra], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    kv = dict(l.split(":", 1) for l in r.stdout.splitlines() if ":" in l and not l.startswith("="))
    assert int(kv["NBFS"]) == 16
    assert float(kv["min_nedge"]) == float(kv["max_nedge"]) == 1048570
    assert float(kv["harmonic_mean_TEPS"]) > 0


def test_driver_logfile_resume_and_tmpfile(tmp_path):
    """LOGFILE: a run keeps its per-root results in the reference's binary
    record and a second run resumes after them (read_log_file /
    update_log_file, mpi/benchmark_helper.hpp:530-604); a record of another
    SCALE is refused.  TMPFILE with --host-edges: the edge list goes through
    a file, which is removed at the end."""
    import struct
    import subprocess
    from graph500_amd import build
    log = tmp_path / "g5.log"
    env = dict(os.environ, LOGFILE=str(log), TMPFILE=str(tmp_path / "edges"))
    say("g5bench 16 LOGFILE, 4 roots")
    r = subprocess.run([build.DRIVER, "16", "--roots", "4", "--host-edges"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert not os.path.exists(str(tmp_path / "edges") + "-000")
    b = log.read_bytes()
    assert len(b) == 1576
    scale, ef, size, runs = struct.unpack_from("<iiii", b, 0)
    assert (scale, ef, size, runs) == (16, 16, 1, 4)
    first = [struct.unpack_from("<ddq", b, 40 + 24 * i) for i in range(4)]
    assert all(t[2] == 1048570 and t[0] > 0 for t in first)
    say("g5bench 16 resume to 8 roots")
    r = subprocess.run([build.DRIVER, "16", "--roots", "8"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "Resume from 4 th run" in r.stdout
    assert "Running BFS 3 ==" not in r.stdout and "========== Running BFS 4 ==========" in r.stdout
    b2 = log.read_bytes()
    assert struct.unpack_from("<i", b2, 12)[0] == 8
    assert [struct.unpack_from("<ddq", b2, 40 + 24 * i) for i in range(4)] == first
    kv = dict(l.split(":", 1) for l in r.stdout.splitlines() if ":" in l and not l.startswith("="))
    assert int(kv["NBFS"]) == 8 and float(kv["max_nedge"]) == 1048570
    say("g5bench 14 with the SCALE 16 record")
    r = subprocess.run([build.DRIVER, "14", "--roots", "2"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 1 and "Log file is not match" in r.stdout


def test_weighted_tuples_on_gpu(oracle_mod):
    """g5_edges_download_weighted: WeightedEdge records whose tuple fields
    equal the packed tuples and whose weights equal the oracle's, in both
    weight forms, across the 8 Mi-tuple staging chunks (2^19 * 17 tuples)."""
    import graph500_amd
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(19, 17, 2, 3)
        m = c.edges_count()
        assert m > 1 << 23
        t = c.edges_download()
        say("weighted download", m, "tuples")
        w0 = c.edges_download_weighted(max_weight=0)
        assert (w0["weight"] == 0xBEEF).all()
        for f in ("v0_low", "v1_low", "high"):
            assert np.array_equal(w0[f], t[f])
        w = c.edges_download_weighted(max_weight=1000)
        for f in ("v0_low", "v1_low", "high"):
            assert np.array_equal(w[f], t[f])
        for lo, hi in [(0, 2000), ((1 << 23) - 1000, (1 << 23) + 1000), (m - 2000, m)]:
            assert np.array_equal(w["weight"][lo:hi], oracle_mod.edge_weights(lo, hi - lo, 1000)), lo
        part = c.edges_download_weighted(start=12345, count=777, max_weight=255, seed1=5, seed2=9)
        assert np.array_equal(part["weight"], oracle_mod.edge_weights(12345, 777, 255, 5, 9))
        with pytest.raises(graph500_amd.G5Error):
            c.edges_download_weighted(start=m - 1, count=2)


@pytest.mark.parametrize("scale", [14, 16])
def test_spec2012_graph_on_gpu(scale, oracle_mod):
    """g5_generate_graph with binary-tree initial tuples (the spec-2012
    generator): the tuple stream, CSR counts and roots equal the reference
    run's, and single-root and batched levels equal the oracle's."""
    import graph500_amd
    g = gold("ref_spec2012_s%d_16roots.json" % scale)
    with graph500_amd.G5Context(0) as c:
        c.generate_graph_spec2012(scale)
        t = c.edges_download().view(np.uint32).reshape(-1, 3)
        assert sha(t) == g["tuples_sha256"]
        c.construct(2)
        info = c.graph_info()
        assert info["num_active"] == g["num_active"] and info["nnz"] == g["nnz"]
        roots = c.find_roots(16)
        assert roots.tolist() == g["roots"]
        oc = oracle_mod.Csr(oracle_mod.generate_graph_spec2012(scale))
        c.prepare_bfs()
        for pr in g["per_root"][:8]:
            c.run_bfs(pr["root"])
            lv_o, _, tr = oc.bfs(pr["root"], want_pred=False)
            assert np.array_equal(c.get_levels(), lv_o), pr["root"]
            st = c.bfs_stats()
            assert "".join("T" if d else "B" for d in st["direction"]) == pr["directions"]
            assert c.teps_edges() == pr["nedge"] == len(t)
        c.batch_prepare()
        c.run_bfs_batch(roots)
        teps = c.batch_teps_edges()
        for i, r in enumerate(roots[:6]):
            lv_o, _, _ = oc.bfs(int(r), want_pred=False)
            assert np.array_equal(c.batch_levels(i), lv_o), r
            assert teps[i] == len(t)
