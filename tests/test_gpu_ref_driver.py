"""The reference's own benchmark driver over the drop-in.

oracle/_ref/runnable_gpu is mpi/main.cc of the reference, compiled from its
source with g5::BfsOnGPU (include/g5_bfs.hpp over libg5.so) in place of
BfsOnCPU (oracle/ref.mk; 64 roots, VERTEX_REORDERING=2, VALIDATION_LEVEL 2).
Everything around the BFS is the reference's code: generate_graph_spec2010
on the host, construct() streaming its EdgeListStorage into HBM, find_roots
over graph_.has_edge, MPI_Wtime around run_bfs, get_pred, its own
validate_bfs_result (mpi/validate.hpp:1008-1019) on every root, and
print_bfs_result.  A run passes when the program exits 0, prints the result
block (it prints "No results printed for invalid run." after a failed
validation), and every root's traversed-edge count is pf_nedge[SCALE]
(mpi/utils.hpp:3681, :3685)."""
import os
import re
import shutil
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "oracle", "_ref", "runnable_gpu")
PF_NEDGE = {16: 1048570, 20: 16776976}


def _mpirun():
    return shutil.which("mpirun") or "/opt/conda/bin/mpirun"


@pytest.mark.parametrize("scale", [16, 20])
def test_reference_main_over_bfs_on_gpu(scale):
    if not os.path.exists(EXE) or not os.path.exists(_mpirun()):
        pytest.skip("oracle/_ref/runnable_gpu or mpirun absent (built from the reference sources by oracle/ref.mk)")
    print("reference main.cc over g5::BfsOnGPU, SCALE %d" % scale, flush=True)
    sys.stdout.flush()
    threads = str(min(16, len(os.sched_getaffinity(0))))
    r = subprocess.run([_mpirun(), "-np", "1", "-genv", "OMP_NUM_THREADS", threads, EXE, str(scale), "16"],
                       capture_output=True, text=True, timeout=300)
    out = r.stdout
    assert r.returncode == 0, out[-3000:] + r.stderr[-2000:]
    assert "No results printed for invalid run." not in out
    kv = dict((a.strip(), b.strip()) for a, b in
              (l.split(":", 1) for l in out.splitlines() if re.match(r"^[a-z_A-Z]+:\s", l)))
    assert int(kv["NBFS"]) == 64
    assert int(kv["SCALE"]) == scale
    assert float(kv["min_nedge"]) == float(kv["max_nedge"]) == PF_NEDGE[scale]
    assert float(kv["harmonic_mean_TEPS"]) > 0
    # one validated count per root, from the reference's own validator (its
    # progress lines go to stderr, the result block to stdout)
    log = out + r.stderr
    counts = [int(x) for x in re.findall(r"Number of traversed edges is (\d+)", log)]
    assert counts == [PF_NEDGE[scale]] * 64
    assert len(re.findall(r"Validate time for BFS \d+ is", log)) == 64
