"""CPU checks of g5bench's log file support (the reference's binary
LogFileFormat, mpi/logfile.h:109-124): --read-log prints a record written in
either byte order as the reference's logreader does, with the statistics of
print_bfs_result (checked against the oracle's restatement,
mpi/logfile.h:19-100).  Resuming a run from such a file needs the GPU and is
tested with the other driver tests."""
import os
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "graph500_amd", "lib", "g5bench")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="g5bench not built")


def _record(endian, scale, ef, runs):
    b = struct.pack(endian + "iiiiddd", scale, ef, 1, len(runs), 1.25, 2.5, 0.0)
    for t in runs:
        b += struct.pack(endian + "ddq", *t)
    return b + b"\0" * (24 * (64 - len(runs)))


def _kv(out):
    return {k.strip(): v.strip() for k, v in (l.split(":", 1) for l in out.splitlines() if ":" in l)}


@pytest.mark.parametrize("endian", ["<", ">"])
def test_read_log_matches_oracle_statistics(tmp_path, oracle_mod, endian):
    rng = np.random.default_rng(9)
    runs = [(float(rng.uniform(0.2, 0.9)), float(rng.uniform(0.01, 0.05)), int(rng.integers(10 ** 6, 2 * 10 ** 6)))
            for _ in range(7)]
    p = tmp_path / "log.bin"
    p.write_bytes(_record(endian, 20, 16, runs))
    assert len(p.read_bytes()) == 1576
    r = subprocess.run([DRIVER, "--read-log", str(p)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "Current Result" in r.stdout  # fewer than 64 runs
    kv = _kv(r.stdout)
    assert int(kv["SCALE"]) == 20 and int(kv["edgefactor"]) == 16
    assert float(kv["graph_generation"]) == 1.25 and float(kv["construction_time"]) == 2.5
    t = np.array([x[0] for x in runs])
    e = np.array([x[2] for x in runs], np.float64)
    v = np.array([x[1] for x in runs])
    for name, arr in (("time", t), ("nedge", e), ("validate", v)):
        st = oracle_mod.statistics(arr)
        for key, so in (("min", "min"), ("firstquartile", "q1"), ("median", "median"), ("thirdquartile", "q3"),
                        ("max", "max"), ("mean", "mean"), ("stddev", "std")):
            assert float(kv["%s_%s" % (key, name)]) == pytest.approx(st[so], rel=1e-9), (key, name)
    hm = 1.0 / np.mean(t / e)
    assert float(kv["harmonic_mean_TEPS"]) == pytest.approx(hm, rel=1e-9)
    lines = r.stdout.splitlines()
    for i, (bt, vt, ec) in enumerate(runs):
        assert "Time for BFS %d is %f" % (i, bt) in lines
        assert "Validate time for BFS %d is %f" % (i, vt) in lines
        assert "TEPS for BFS %d is %g" % (i, ec / bt) in lines


def test_read_log_missing_file(tmp_path):
    r = subprocess.run([DRIVER, "--read-log", str(tmp_path / "none.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "File not found" in r.stderr
