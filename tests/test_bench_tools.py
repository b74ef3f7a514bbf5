"""CPU checks of bench.py's contract and of its rocprof helpers: the
multi-GPU refusal, the traffic figure read from the committed profile, the
request-rate ceiling read from the committed gather calibration, and
the summaries of a rocprofv3 results database (a small stand-in database with
the tables and columns the helpers read)."""
import os
import sqlite3
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one MI355X" in r.stderr


def test_traffic_from_committed_profile():
    t = bench.traffic_from_profile("k_ms_mixed", 28)
    assert t is not None and t > 1e10
    text = open(os.path.join(ROOT, "profiles", "k_ms_mixed_s28_traffic.txt")).read()
    assert "traffic_bytes_per_launch: %.0f" % t in text
    assert bench.traffic_from_profile("no_such_kernel", 28) is None


def test_request_peak_from_committed_calibration():
    """The request-rate ceiling is the best random-read rate of the committed
    gather calibration over arrays past the L2s (32 MiB and up), read from
    its table; restricted to 1 GiB and up it is the 1 GiB row."""
    text = open(os.path.join(ROOT, "profiles", "random_gather_rate.txt")).read()
    peak = bench.gather_peak_beyond_l2()
    assert peak is not None and 40.0 < peak < 100.0
    assert " %.1f " % peak in text
    big = bench.gather_peak_beyond_l2(1 << 30)
    assert big is not None and big <= peak
    assert bench.gather_peak_beyond_l2(1 << 40) is None


def _trace_db(path):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, start integer, end integer)")
    con.execute("create table top_kernels (name text, total_calls integer, total_duration real, average real, "
                "percentage real)")
    rows = [("g5::k_ms_pull_lanes(unsigned long*)", 0, 26_000_000), ("g5::k_ms_push_list(unsigned int const*)",
            27_000_000, 29_000_000), ("g5::k_ms_pull_lanes(unsigned long*)", 30_000_000, 31_500_000)]
    con.executemany("insert into kernels values (?, ?, ?)", rows)
    con.execute("insert into top_kernels values ('g5::k_ms_pull_lanes(unsigned long*)', 2, 27500.0, 13750.0, "
                "93.2)")
    con.commit()
    con.close()


def _pmc_db(path):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (dispatch_id integer, kernel_name text, counter_name text, "
                "value real)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)",
                    [(2, "g5::k_ms_pull_lanes(unsigned long*)", "FETCH_SIZE", 300.0),
                     (1, "g5::k_ms_pull_lanes(unsigned long*)", "FETCH_SIZE", 100.0),
                     (3, "g5::k_ms_select(unsigned long const*)", "FETCH_SIZE", 7.0)])
    con.commit()
    con.close()


def test_rocprof_dispatches_in_order(tmp_path):
    db = str(tmp_path / "trace.db")
    _trace_db(db)
    lines = bench.rocprof_dispatches(db, "k_ms_pull_lanes").splitlines()
    assert [float(l.split()[-1]) for l in lines] == [26.0, 1.5]
    db2 = str(tmp_path / "pmc.db")
    _pmc_db(db2)
    vals = [float(l.split()[-1]) for l in bench.rocprof_dispatches(db2, "k_ms_pull_lanes").splitlines()]
    assert vals == [100.0, 300.0]  # dispatch order, not insertion order


def test_rocprof_summary_both_kinds(tmp_path):
    db = str(tmp_path / "trace.db")
    _trace_db(db)
    s = bench.rocprof_summary(db)
    assert "g5::k_ms_pull_lanes" in s and "27.500" in s  # total ms from the view's microseconds
    db2 = str(tmp_path / "pmc.db")
    _pmc_db(db2)
    s2 = bench.rocprof_summary(db2)
    line = [l for l in s2.splitlines() if "k_ms_pull_lanes" in l][0]
    assert line.split()[-2] == "2" and float(line.split()[-1]) == pytest.approx(200.0)


def test_rocprof_overlap_by_brute_force(tmp_path):
    """Busy time and the time with 2+ / 3+ kernels at once equal a count over
    every nanosecond of a small random trace."""
    import numpy as np
    rng = np.random.default_rng(3)
    unit = 10_000  # ns: the trace spans about 24 ms
    rows = []
    for i in range(40):
        s = int(rng.integers(0, 2000))
        rows.append(("g5::k_%s(int)" % ("sv_pull" if i else "build"), s, s + int(rng.integers(1, 300)),
                     int(rng.integers(0, 5))))
    rows.sort(key=lambda r: r[1])
    rows[0] = ("g5::k_build(int)",) + rows[0][1:]
    rows[1:] = [("g5::k_sv_pull(int)",) + r[1:] for r in rows[1:]]
    scaled = [(n, s * unit, e * unit, st) for n, s, e, st in rows]
    db = str(tmp_path / "trace.db")
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, start integer, end integer, stream_id integer)")
    con.executemany("insert into kernels values (?, ?, ?, ?)", scaled)
    con.commit()
    con.close()
    cover = np.zeros(2400, np.int64)
    for _, s, e, _ in rows[1:]:  # from the first k_sv_ dispatch on
        cover[s:e] += 1
    ms = unit / 1e6
    out = bench.rocprof_overlap(db, "k_sv_").splitlines()
    busy = float(out[1].split(": ")[1].split()[0])
    two = float(out[2].split(": ")[1].split()[0])
    three = float(out[3].split(": ")[1].split()[0])
    assert (cover > 2).sum() > 0  # the trace has 3-deep overlaps
    assert busy == pytest.approx((cover > 0).sum() * ms, abs=0.06)
    assert two == pytest.approx((cover > 1).sum() * ms, abs=0.06)
    assert three == pytest.approx((cover > 2).sum() * ms, abs=0.06)
    assert int(out[4].split(": ")[1]) == cover.max()
    assert bench.rocprof_overlap(db, "k_nothing").startswith("no dispatch")
