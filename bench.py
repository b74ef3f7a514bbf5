#!/usr/bin/env python3
"""Graph500 BFS hot path on one MI355X at Kronecker SCALE 28, edgefactor 16
(seeds 2,3; initiator 0.57/0.19/0.19/0.05): 64-root batched BFS against
single-root BFS, GTEPS.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--scale 28]

A step is one 64-root batched BFS (g5_run_bfs_batch) over the graph resident
in HBM, for the 64 roots find_roots draws (mpi/benchmark_helper.hpp:475-508):
one edge sweep per level serves all 64 searches, each root's levels bit-exact
with its single-root run.  `value` = sum over roots of traversed edges / batch
time (host clock around the synchronous call), the batched analogue of the
harmonic mean (mpi/logfile.h:74-82).  Generation and CSR construction run on
the GPU before the timed region, as in the reference driver (mpi/main.cc:
50-117, only run_bfs is timed).

Before the timed steps the autotuner (graph500_amd/autotune.py) picks the
batch's launch shapes and switch thresholds on this graph and these roots.

`single_root` times the drop-in path, g5_run_bfs, once per root (after
warm-up), with the host clock around each synchronous call like MPI_Wtime
around run_bfs; GTEPS = 1 / mean(t_i / e_i) with e_i the validator's
traversed-edge count.

`roofline` is for the dominant kernel of the step (the kind of batch level
with the most device time: push, vertex-parallel pull, edge-parallel sweep,
or a level where some roots push and the others pull): algorithmic bytes of its launches (byte model evaluated on the
GPU from the final levels, outside timing; equal to the CPU oracle's count,
see tests) over their HIP-event durations on the library stream.  `traffic`
is the HBM bytes per launch measured by rocprof (FETCH_SIZE + WRITE_SIZE,
scaled as the plain-copy and random-gather calibrations of the same counters
direct) when a summary for this kernel and workload is present under
profiles/, else null; `request_frac` is the kernel's memory-request rate
over the best random-read rate past the L2s in the gather calibration.
`serve` quotes its policy: short queues on single-root slots, longer ones
in lanes.  `cpu_baseline` times
the reference CPU runnable (oracle/_ref/runnable) on a bounded sample on this
host, or the oracle port when that cannot run.

`serve` runs the serving loop (g5_server_*) over resident graphs of SCALE
18, 22, 25 and 27 at swept Poisson rates: p50/p99 latency from arrival to a
complete answer, sustained GTEPS, and a sample of answers checked against
single-root runs.

Other modes:
  --copy-calibration      a plain 1 GiB device copy (for scaling rocprof
                          FETCH_SIZE/WRITE_SIZE counters)
  --gather-calibration    2^28 random 8-byte reads over a 4 GiB array (what the
                          same counters read per random 8-byte miss)
  --rocprof-summary DB    per-kernel time, or per-kernel counter averages, of a
                          rocprofv3 results database (the text kept in profiles/)
  --rocprof-dispatches DB KERNEL
                          per-dispatch durations or counters of one kernel
  --rocprof-overlap DB KERNEL
                          how kernels overlap on the device from the first
                          dispatch of KERNEL on (busy time, time with 2+ and 3+
                          kernels at once, dispatches per stream)
  --tuning JSON           run the batch with this tuning instead of autotuning
  --single-root-only N    only the single-root leg, over the first N roots (a
                          short run to profile under rocprofv3)
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "GTEPS hmean SCALE 28 ef16: 64-root batch vs single-root; serve p50/p99 ms"
HBM_PEAK_GBS = 8000.0
STATED_LOAD_QPS = 1000.0  # the serving load at which p50/p99 are quoted
# batch level kinds (g5_batch_stats.direction) -> (profile name, kernel description)
BATCH_KERNELS = {
    0: ("k_ms_pull", "k_ms_pull (batched vertex-parallel pull)"),
    1: ("k_ms_push", "k_ms_push + k_ms_finish (batched push)"),
    2: ("k_ms_sweep", "k_ms_sweep + k_ms_finish (batched edge-parallel pull sweep)"),
    3: ("k_ms_mixed", "k_ms_select + queue scans + k_chunk_rows + k_ms_push + k_ms_pull_lanes (batched level "
                      "with per-root directions: the pushing roots' frontier as a queue, then the pulling roots)"),
}
# the mixed level's kernels when its push goes over CSR chunks (mixed_push 0)
MIXED_CHUNK_PUSH = ("k_ms_select + k_ms_collect + k_ms_push_list + k_ms_pull_lanes (batched level with per-root "
                    "directions: the pushing roots over CSR chunks, then the pulling roots)")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    """Host facts the CPU baseline is quoted with: the CPU model, nproc, and
    the cores this process may use (its affinity mask)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0))}


def baseline_threads():
    """All cores of this process's share: its affinity mask, capped by
    OMP_NUM_THREADS where the host sets one (a GPU box gives each GPU's jobs
    a share of its cores and sets OMP_NUM_THREADS to it)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(threads, scale=20, timeout=150):
    """Reference runnable at a bounded SCALE; falls back to the oracle port."""
    exe = os.path.join(HERE, "oracle", "_ref", "runnable")
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    host = cpu_info()
    if os.path.exists(exe) and os.path.exists(mpirun):
        try:
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            t0 = time.time()
            r = subprocess.run([mpirun, "-np", "1", "-genv", "OMP_NUM_THREADS", str(threads), exe,
                                str(scale), "16"], capture_output=True, text=True, timeout=timeout,
                               env=env)
            m = re.search(r"harmonic_mean_TEPS:\s+([0-9.eE+-]+)", r.stdout)
            if r.returncode == 0 and m:
                keys = {}
                for k in ("mean_time", "median_TEPS", "construction_time", "graph_generation", "mean_validate"):
                    mk = re.search(r"^%s:\s+([0-9.eE+-]+)" % k, r.stdout, re.M)
                    keys[k] = float(mk.group(1)) if mk else None
                out = {
                    "value": float(m.group(1)) / 1e9,
                    "unit": "GTEPS",
                    "cores": threads,
                    "kind": "reference",
                    "sample": ("reference runnable (VERTEX_REORDERING=2, 64 roots), mpirun -np 1, "
                               "OMP_NUM_THREADS=%d (every core of this job's share), Kronecker SCALE %d ef16; "
                               "harmonic-mean TEPS of its own timed run_bfs calls (mpi/main.cc:113-117); "
                               "wall %.1f s including generation, construction and validation. SCALE %d keeps "
                               "the sample within the bench's budget: the 1-process reference aborts beyond "
                               "SCALE 24 and needs minutes there" % (threads, scale, time.time() - t0, scale)),
                    "scale": scale,
                }
                out.update(keys)
                out.update(host)
                return out
            log("reference runnable failed rc=%s: %s" % (r.returncode, r.stderr[-400:]))
        except Exception as ex:  # noqa: BLE001
            log("reference runnable unavailable: %r" % (ex,))
    # oracle port, single thread
    from oracle import oracle as O
    O.build()
    e = O.generate_edges(scale)
    c = O.Csr(e)
    roots = c.find_roots(16)
    spe = []
    for r in roots:
        t0 = time.perf_counter()
        _, _, tr = c.bfs(int(r), want_pred=True)
        spe.append((time.perf_counter() - t0) / tr["teps_edges"])
    out = {"value": len(spe) / sum(spe) / 1e9, "unit": "GTEPS", "cores": 1, "kind": "port",
           "sample": "oracle BFS (single thread) over 16 roots at SCALE %d ef16" % scale, "scale": scale}
    out.update(host)
    return out


def copy_calibration():
    import torch
    n = 1 << 28  # 2^28 int32 = 1 GiB
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))


def gather_calibration(sizes=(4 << 30,), gathers=1 << 28):
    """One launch of k_measure_gather (libg5.so, include/g5_measure.h) per
    array size: `gathers` random 8-byte reads over an array of that many
    bytes.  Over an array far larger than the last-level cache nearly every
    read misses; read under rocprofv3 --pmc, it tells what the traffic
    counters report per random 8-byte miss.  Over a range of sizes it gives
    the random-read rate of each level of the memory system (L2-resident,
    last-level-cache-resident, HBM)."""
    import ctypes
    from graph500_amd import g5
    lib = g5.load_library()
    for array_bytes in sizes:
        ms = ctypes.c_double(0)
        chk = ctypes.c_uint64(0)
        rc = lib.g5_measure_gather(0, ctypes.c_int64(array_bytes), ctypes.c_int64(gathers), ctypes.byref(ms),
                                   ctypes.byref(chk))
        if rc != 0:
            raise RuntimeError("g5_measure_gather failed: %d" % rc)
        rate = gathers / (ms.value / 1e3)
        print(json.dumps({"kernel": "k_measure_gather", "array_bytes": array_bytes, "gathers": gathers,
                          "ms": round(ms.value, 4), "gathers_per_s": rate,
                          "gbs_if_64B_per_gather": round(rate * 64 / 1e9, 1),
                          "gbs_if_128B_per_gather": round(rate * 128 / 1e9, 1), "checksum": chk.value}),
              flush=True)


def rocprof_summary(db, top=30):
    """Text summary of a rocprofv3 SQLite results database: per-kernel calls,
    total and average duration from a --kernel-trace run, or per-kernel
    dispatch-averaged counter values from a --pmc run."""
    import collections
    import sqlite3
    cur = sqlite3.connect(db).cursor()
    tables = {r[0] for r in cur.execute("select name from sqlite_master where type in ('table','view')")}
    lines = []
    if "counters_collection" in tables and cur.execute("select count(*) from counters_collection").fetchone()[0]:
        agg = collections.defaultdict(lambda: [0, 0.0])
        for k, c, v in cur.execute("select kernel_name, counter_name, value from counters_collection"):
            key = (k.split("(")[0], c)
            agg[key][0] += 1
            agg[key][1] += v
        lines.append("%-48s %-12s %10s %18s" % ("kernel", "counter", "dispatches", "avg per dispatch"))
        for (k, c), (n, tot) in sorted(agg.items(), key=lambda x: -x[1][1])[:top]:
            lines.append("%-48s %-12s %10d %18.1f" % (k[:48], c, n, tot / n))
    else:
        rows = cur.execute("select name, total_calls, total_duration, average, percentage "
                           "from top_kernels").fetchall()
        lines.append("%-48s %8s %12s %12s %7s" % ("kernel", "calls", "total ms", "avg us", "pct"))
        for name, calls, tot, avg, pct in rows[:top]:  # the view's durations are in microseconds
            lines.append("%-48s %8d %12.3f %12.2f %7.2f" % (name.split("(")[0][:48], calls, tot / 1e3, avg, pct))
    return "\n".join(lines)


def rocprof_dispatches(db, pattern):
    """Per-dispatch values, in dispatch order, of the kernels whose name
    contains `pattern`: duration (ms) from a --kernel-trace database, or each
    counter from a --pmc database."""
    import sqlite3
    cur = sqlite3.connect(db).cursor()
    tables = {r[0] for r in cur.execute("select name from sqlite_master where type in ('table','view')")}
    lines = []
    if "counters_collection" in tables and cur.execute("select count(*) from counters_collection").fetchone()[0]:
        cols = {r[1] for r in cur.execute("pragma table_info(counters_collection)")}
        order = "dispatch_id" if "dispatch_id" in cols else "rowid"
        for k, c, v in cur.execute("select kernel_name, counter_name, value from counters_collection "
                                   "order by %s" % order):
            if pattern in k:
                lines.append("%-40s %-12s %.1f" % (k.split("(")[0][:40], c, v))
    else:
        for k, s, e in cur.execute("select name, start, end from kernels order by start"):
            if pattern in k:
                lines.append("%-40s %.4f" % (k.split("(")[0][:40], (e - s) / 1e6))
    return "\n".join(lines)


def rocprof_overlap(db, pattern):
    """How the kernels of a --kernel-trace database overlap on the device,
    from the first dispatch whose name contains `pattern` to the last
    dispatch: busy time (the union of kernel intervals), the time with two
    or more and with three or more kernels running at once, the largest
    number at once, and the dispatches per stream."""
    import collections
    import sqlite3
    cur = sqlite3.connect(db).cursor()
    rows = list(cur.execute("select name, start, end, stream_id from kernels order by start"))
    first = next((i for i, r in enumerate(rows) if pattern in r[0]), None)
    if first is None:
        return "no dispatch matches %r" % pattern
    rows = rows[first:]
    ev = sorted([(r[1], 1) for r in rows] + [(r[2], -1) for r in rows], key=lambda x: (x[0], x[1]))
    depth, last, top = 0, ev[0][0], 0
    t = collections.Counter()
    for when, d in ev:
        if depth > 0:
            t[min(depth, 3)] += when - last
        depth += d
        top = max(top, depth)
        last = when
    busy = sum(t.values())
    window = rows[-1][2] - rows[0][1] if rows else 0
    per_stream = collections.Counter(r[3] for r in rows)
    lines = ["window from the first %s dispatch: %.1f ms, %d dispatches on %d streams" %
             (pattern, window / 1e6, len(rows), len(per_stream)),
             "device busy (some kernel running): %.1f ms (%.1f %% of the window)" % (busy / 1e6, 100.0 * busy / max(window, 1)),
             "two or more kernels at once: %.1f ms (%.1f %% of busy)" % ((t[2] + t[3]) / 1e6, 100.0 * (t[2] + t[3]) / max(busy, 1)),
             "three or more kernels at once: %.1f ms (%.1f %% of busy)" % (t[3] / 1e6, 100.0 * t[3] / max(busy, 1)),
             "most kernels at once: %d" % top,
             "dispatches per stream: %s" % ", ".join("%d" % c for _, c in per_stream.most_common())]
    return "\n".join(lines)


def traffic_from_profile(kernel, scale):
    """HBM bytes per launch of `kernel` from the committed rocprof PMC summary
    (profiles/<kernel>_s<scale>_traffic.txt, line 'traffic_bytes_per_launch: X')."""
    return profile_figure(kernel, scale, "traffic_bytes_per_launch")


def profile_figure(kernel, scale, key):
    """A 'key: value' figure of the committed profile summary of `kernel` at
    `scale` (profiles/<kernel>_s<scale>_traffic.txt), or None."""
    path = os.path.join(HERE, "profiles", "%s_s%d_traffic.txt" % (kernel, scale))
    if not os.path.exists(path):
        return None
    m = re.search(r"^%s:\s*([0-9.eE+]+)" % re.escape(key), open(path).read(), re.M)
    return float(m.group(1)) if m else None


def gather_peak_beyond_l2(min_bytes=32 << 20):
    """The best random 8-byte read rate (G reads/s) of the committed gather
    calibration (profiles/random_gather_rate.txt) over arrays of at least
    min_bytes, i.e. larger than the 8 XCD L2s together: the request-rate
    ceiling of random reads that miss the L2, or None."""
    path = os.path.join(HERE, "profiles", "random_gather_rate.txt")
    if not os.path.exists(path):
        return None
    unit = {"KiB": 1 << 10, "MiB": 1 << 20, "GiB": 1 << 30}
    best = None
    for m in re.finditer(r"^\s*(\d+) (KiB|MiB|GiB)\s+([0-9.]+)\s+([0-9.]+)", open(path).read(), re.M):
        if int(m.group(1)) * unit[m.group(2)] >= min_bytes:
            rate = float(m.group(4))
            best = rate if best is None else max(best, rate)
    return best


def single_root_leg(ctx, roots, warmup):
    """Each root once through g5_run_bfs; host clock per call."""
    ctx.prepare_bfs()
    for i in range(warmup):
        ctx.run_bfs(int(roots[i % len(roots)]))
    ctx.synchronize()
    wall, dev, edges, tot_bytes, bu_bytes, bu_ms, bu_launches = [], [], [], 0.0, 0.0, 0.0, 0
    hub_bytes, hub_ms, hub_levels = 0.0, 0.0, 0
    for i, r in enumerate(roots):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run_bfs(int(r))
        t1 = time.perf_counter()
        st = ctx.bfs_stats()
        e = ctx.teps_edges()
        per, tot = ctx.model_bytes()
        wall.append(t1 - t0)
        dev.append(st["device_ms"] / 1e3)
        edges.append(e)
        tot_bytes += tot
        for L, d in enumerate(st["direction"]):
            if d == 0:
                bu_bytes += per[L]
                bu_ms += st["level_ms"][L]
                bu_launches += 1
            elif not st["bu_kernel"][L] and hub_edges(per[L], st["nq"], L) > (1 << 20):
                # a bulk top-down level (more than 2^20 frontier edges)
                hub_bytes += per[L]
                hub_ms += st["level_ms"][L]
                hub_levels += 1
        if i % 16 == 0:
            log("single root %d: %.3f ms %.1f GTEPS %s" % (
                r, 1e3 * (t1 - t0), e / (t1 - t0) / 1e9,
                "".join("T" if d else "B" for d in st["direction"])))
    ctx.end_bfs()
    hmean = len(wall) / sum(w / e for w, e in zip(wall, edges)) / 1e9
    hmean_dev = len(dev) / sum(w / e for w, e in zip(dev, edges)) / 1e9
    bu_gbs = bu_bytes / (bu_ms / 1e3) / 1e9 if bu_ms > 0 else 0.0
    hub_gbs = hub_bytes / (hub_ms / 1e3) / 1e9 if hub_ms > 0 else 0.0
    return {
        "gteps_hmean_wall": round(hmean, 3),
        "gteps_hmean_device": round(hmean_dev, 3),
        "roots": len(roots),
        "mean_wall_ms": round(1e3 * sum(wall) / len(wall), 4),
        "mean_device_ms": round(1e3 * sum(dev) / len(dev), 4),
        "bytes_per_edge": round(tot_bytes / sum(edges), 4),
        "roofline_frac_whole_bfs": round((tot_bytes / sum(dev)) / (HBM_PEAK_GBS * 1e9), 4),
        "traversed_edges_per_root": int(edges[0]),
        "k_bu": {"achieved_gbs": round(bu_gbs, 1), "frac": round(bu_gbs / HBM_PEAK_GBS, 4),
                 "launches": bu_launches,
                 "algorithmic_bytes_per_launch": round(bu_bytes / max(bu_launches, 1)),
                 "avg_launch_ms": round(bu_ms / max(bu_launches, 1), 4)},
        # bulk top-down levels: head (k_td_head), row suffixes (k_td_expand_segs)
        # or rows (k_td_expand), and k_td_finish, per level
        "bulk_td_levels": {"achieved_gbs": round(hub_gbs, 1), "frac": round(hub_gbs / HBM_PEAK_GBS, 4),
                           "levels": hub_levels,
                           "algorithmic_bytes_per_level": round(hub_bytes / max(hub_levels, 1)),
                           "avg_level_ms": round(hub_ms / max(hub_levels, 1), 4)},
    }


def hub_edges(level_bytes, nq, L):
    """Frontier edges E of top-down level L + 1 (list index L), from its
    byte-model figure: 12 B per frontier vertex + 4 B per frontier edge + 5 B
    per discovered vertex (level byte and queue entry)."""
    nf = nq[L - 1] if L > 0 else 1
    return (level_bytes - 5 * nq[L] - 12 * nf) / 4


# Serving policy (continuous batching): a graph with searches in flight admits
# waiting queries into its free lanes at a level boundary while every search
# in flight still pushes, once min(min_admit, free answer arrays) queries can
# start together, or once the oldest has waited max_wait_s; each search pulls
# while its frontier's edges exceed 1/alpha of its unreached edges.  A graph
# with no search in lanes and at most 8 waiting queries answers them with the
# single-root kernels, up to 3 at once on slots of their own (solo,
# solo_slots); a longer queue is batched into lanes.
SERVE_POLICY = {"answer_buffers": 128, "min_admit": 64, "max_wait_s": 0.3, "lane_rule": 3, "alpha": 64.0,
                "solo": 8, "solo_slots": 3}


def serve_leg(scales, rates, max_queries, edgefactor, policy=None, seed=1):
    """Serving loop over resident graphs: Poisson arrivals at each rate, each
    query on graph s with probability proportional to 2^-((27 - s) / 2) and a
    uniform random root with edges.  Per rate: latency p50/p99 (arrival to
    complete answer), sustained GTEPS (traversed edges / wall), and a sample
    of kept answers compared with single-root runs of the same roots."""
    import numpy as np
    import graph500_amd
    ctxs = []
    for sc in scales:
        c = graph500_amd.G5Context(0)
        c.generate_kronecker(sc, edgefactor, 2, 3)
        c.construct(2, release_edges=True)
        c.prepare_bfs()
        ctxs.append(c)
        log("serve graph SCALE %d: n_active %d" % (sc, c.graph_info()["num_active"]))
    w = np.array([2.0 ** (-(27 - sc) / 2.0) for sc in scales])
    w /= w.sum()
    rng = np.random.default_rng(seed)
    has = []
    for c in ctxs:
        n = c.pred_size()
        cand = rng.integers(0, n, size=min(n, 200000))
        has.append(np.array([v for v in cand[:20000] if c.has_edge(int(v))], np.int64))
    runs = []
    policy = dict(SERVE_POLICY, **(policy or {}))
    with graph500_amd.G5Server(ctxs, **policy) as srv:
        for rate in rates:
            nq = int(min(max_queries, max(64, rate * 1.0)))
            q = np.zeros(nq, graph500_amd.QUERY_DTYPE)
            q["graph"] = rng.choice(len(scales), size=nq, p=w)
            q["root"] = [has[g][rng.integers(0, len(has[g]))] for g in q["graph"]]
            q["arrival_s"] = np.cumsum(rng.exponential(1.0 / rate, size=nq))
            check = rng.choice(nq, size=min(4, nq), replace=False)
            q["keep"][check] = 1
            solo0 = sum(srv.graph_stats(g)["solo"] for g in range(len(scales)))
            over0 = sum(srv.graph_stats(g)["solo_overlap"] for g in range(len(scales)))
            res, rep = srv.run(q, timeout_s=300)
            solo = sum(srv.graph_stats(g)["solo"] for g in range(len(scales))) - solo0
            over = sum(srv.graph_stats(g)["solo_overlap"] for g in range(len(scales))) - over0
            lat = (res["done_s"] - q["arrival_s"]) * 1e3
            ok = True
            for i in check:
                g, r = int(q["graph"][i]), int(q["root"][i])
                lv = srv.levels(res["request"][i], ctxs[g].pred_size())
                srv.release(res["request"][i])
                ctxs[g].run_bfs(r)
                ok &= bool(np.array_equal(lv, ctxs[g].get_levels())) and int(res["edges"][i]) == ctxs[g].teps_edges()
            per_graph = {}
            for gi, sc in enumerate(scales):
                lg = lat[q["graph"] == gi]
                if len(lg):
                    per_graph["S%d" % sc] = {"queries": int(len(lg)), "p50_ms": round(float(np.percentile(lg, 50)), 3),
                                             "p99_ms": round(float(np.percentile(lg, 99)), 3)}
            run = {"rate_qps": rate, "queries": nq, "p50_ms": round(float(np.percentile(lat, 50)), 3),
                   "p99_ms": round(float(np.percentile(lat, 99)), 3), "per_graph": per_graph,
                   "served_qps": round(nq / rep["wall_s"], 1),
                   "gteps": round(rep["edges"] / rep["wall_s"] / 1e9, 2),
                   "max_lanes": rep["max_active_lanes"], "admitted_busy": rep["admitted_busy"],
                   "admit_events": rep["admit_events"], "single_root_answers": int(solo),
                   "single_root_overlapped": int(over),
                   "sample_bit_exact": ok}
            log("serve %s" % run)
            runs.append(run)
    for c in ctxs:
        c.close()
    stated = next((r for r in runs if r["rate_qps"] == STATED_LOAD_QPS), None)
    return {"graphs_scale": list(scales), "policy": policy, "stated_load_qps": STATED_LOAD_QPS,
            "p50_ms_at_stated_load": stated["p50_ms"] if stated else None,
            "p99_ms_at_stated_load": stated["p99_ms"] if stated else None,
            "admitted_busy_at_stated_load": stated["admitted_busy"] if stated else None,
            "single_root_overlapped_at_stated_load": stated["single_root_overlapped"] if stated else None,
            "runs": runs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=28)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core of this job's share")
    ap.add_argument("--cpu-scale", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-single-root", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-scales", type=lambda x: [int(v) for v in x.split(",")], default=[18, 22, 25, 27])
    ap.add_argument("--serve-rates", type=lambda x: [float(v) for v in x.split(",")],
                    default=[100.0, 300.0, 1000.0, 3000.0, 20000.0])
    ap.add_argument("--serve-max-queries", type=int, default=3000)
    ap.add_argument("--serve-policy", type=json.loads, default=None,
                    help="serving policy overrides as JSON, e.g. {\"min_admit\": 16}")
    ap.add_argument("--serve-only", action="store_true", help="only the serving leg (no SCALE 28 graph)")
    ap.add_argument("--autotune-budget", type=float, default=30.0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--gather-calibration", action="store_true")
    ap.add_argument("--gather-bytes", type=lambda x: [int(v) for v in x.split(",")], default=[4 << 30],
                    help="array sizes of --gather-calibration, bytes, comma-separated")
    ap.add_argument("--rocprof-summary", metavar="DB")
    ap.add_argument("--rocprof-dispatches", nargs=2, metavar=("DB", "KERNEL"))
    ap.add_argument("--rocprof-overlap", nargs=2, metavar=("DB", "KERNEL"))
    ap.add_argument("--single-root-only", type=int, metavar="N")
    ap.add_argument("--tuning", type=json.loads, default=None,
                    help="batch tuning as JSON (skips the autotuner), e.g. a previous run's batch.tuning")
    a = ap.parse_args()

    if a.rocprof_summary:
        print(rocprof_summary(a.rocprof_summary))
        return 0
    if a.rocprof_dispatches:
        print(rocprof_dispatches(*a.rocprof_dispatches))
        return 0
    if a.rocprof_overlap:
        print(rocprof_overlap(*a.rocprof_overlap))
        return 0
    if a.copy_calibration:
        copy_calibration()
        return 0
    if a.gather_calibration:
        gather_calibration(a.gather_bytes)
        return 0
    if a.gpus != 1:
        print("bench.py measures one MI355X only: this build has no multi-GPU path "
              "(--gpus %d requested)" % a.gpus, file=sys.stderr)
        return 2

    if a.serve_only:
        print(json.dumps({"mode": "serve only", "serve": serve_leg(a.serve_scales, a.serve_rates,
                                                                    a.serve_max_queries, a.edgefactor,
                                                                    a.serve_policy)}))
        return 0

    import graph500_amd
    ctx = graph500_amd.G5Context(0)
    log("generate SCALE %d ef %d" % (a.scale, a.edgefactor))
    ctx.generate_kronecker(a.scale, a.edgefactor, 2, 3)
    log("construct")
    ctx.construct(2, release_edges=True)
    info = ctx.graph_info()
    log("graph: n_active %d nnz %d gen %.1f ms construct %.1f ms" %
        (info["num_active"], info["nnz"], info["generate_ms"], info["construct_ms"]))
    roots = ctx.find_roots(64)

    if a.single_root_only:
        single = single_root_leg(ctx, roots[:a.single_root_only], a.warmup)
        ctx.close()
        print(json.dumps({"mode": "single-root only", "scale": a.scale, "single_root": single}))
        return 0

    single = None
    if not a.no_single_root:
        single = single_root_leg(ctx, roots, a.warmup)
        log("single root: %.1f GTEPS hmean" % single["gteps_hmean_wall"])

    ctx.batch_prepare()
    tuned = None
    if a.tuning:
        ctx.batch_set_tuning(**a.tuning)
    elif not a.no_autotune:
        from graph500_amd import autotune
        tuned = autotune.autotune(ctx, roots, reps=3, budget_s=a.autotune_budget, use_cache=False, log=log)
        log("autotune: %s -> %.3f ms (%d trials, %.1f s)" % (tuned["tuning"], tuned["ms"], tuned["trials"],
                                                             tuned["seconds"]))
    for _ in range(a.warmup):
        ctx.run_bfs_batch(roots)
    ctx.synchronize()
    edges = ctx.batch_teps_edges()
    total_edges = int(edges.sum())
    per, tot = ctx.batch_model_bytes()  # the schedule is the same every step
    wall, dev = [], []
    # per level kind (push / vertex-parallel pull / edge-parallel sweep): launches, ms, model bytes
    kinds = {d: [0, 0.0, 0.0] for d in BATCH_KERNELS}
    t_all = time.perf_counter()
    for i in range(a.steps):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run_bfs_batch(roots)
        t1 = time.perf_counter()
        st = ctx.batch_stats()
        wall.append(t1 - t0)
        dev.append(st["device_ms"] / 1e3)
        for L, d in enumerate(st["direction"]):
            kinds[d][0] += 1
            kinds[d][1] += st["level_ms"][L]
            kinds[d][2] += per[L]
        log("batch step %d: %.3f ms %.1f GTEPS" % (i, 1e3 * (t1 - t0), total_edges / (t1 - t0) / 1e9))
    t_all = time.perf_counter() - t_all
    st = ctx.batch_stats()
    assert (ctx.batch_teps_edges() == edges).all()
    mean_wall = sum(wall) / len(wall)
    gteps = total_edges / mean_wall / 1e9
    dom = max(kinds, key=lambda d: kinds[d][1])  # the level kind with the most device time
    dom_launches, dom_ms, dom_bytes = kinds[dom]
    dom_gbs = dom_bytes / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    per_launch = dom_bytes / max(dom_launches, 1)
    traffic = traffic_from_profile(BATCH_KERNELS[dom][0], a.scale)
    requests = profile_figure(BATCH_KERNELS[dom][0], a.scale, "memory_requests_per_launch")
    gather_rate = profile_figure(BATCH_KERNELS[dom][0], a.scale, "random_read_rate_g_per_s")
    request_peak = gather_peak_beyond_l2()
    ctx_tuning = ctx.batch_tuning()
    out = {
        "metric": METRIC,
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * mean_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: Kronecker graph from the Graph500 splittable-MRG generator "
                "(seeds 2,3), generated and built on the GPU",
        "config": {
            "workload": "64-root batched BFS (one bit per root in u64 per-vertex words), Kronecker "
                        "SCALE %d edgefactor %d, the 64 find_roots roots, VERTEX_REORDERING=2"
                        % (a.scale, a.edgefactor),
            "scale": a.scale,
            "edgefactor": a.edgefactor,
            "num_active": info["num_active"],
            "directed_entries": info["nnz"],
            "roots": len(roots),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": (MIXED_CHUNK_PUSH if dom == 3 and ctx_tuning.get("mixed_push") == 0
                       else BATCH_KERNELS[dom][1]),
            "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "launches": dom_launches,
            "algorithmic_bytes_per_launch": round(per_launch),
            "avg_launch_ms": round(dom_ms / max(dom_launches, 1), 4),
            # the same kernels as memory requests (rocprof TCC_EA0_RDREQ + WRREQ
            # per launch) against the random-read rate of the memory system
            # (k_measure_gather over an array larger than the last-level cache)
            "memory_requests_per_launch": requests,
            "request_rate_g_per_s": (round(requests / (dom_ms / max(dom_launches, 1) / 1e3) / 1e9, 1)
                                     if requests and dom_ms > 0 else None),
            "random_read_rate_g_per_s": gather_rate,
            # the request-rate roofline: the level's requests per second
            # against the best random-read rate beyond the L2s
            "request_peak_g_per_s": request_peak,
            "request_frac": (round(requests / (dom_ms / max(dom_launches, 1) / 1e3) / 1e9 / request_peak, 3)
                             if requests and dom_ms > 0 and request_peak else None),
        },
        "batch": {
            "gteps": round(gteps, 3),
            "gteps_device": round(total_edges / (sum(dev) / len(dev)) / 1e9, 3),
            "traversed_edges_sum": total_edges,
            "bytes_per_edge": round(tot / total_edges, 4),
            "roofline_frac_whole_batch": round((tot / (sum(dev) / len(dev))) / (HBM_PEAK_GBS * 1e9), 4),
            "levels": "".join("UPSM"[d] for d in st["direction"]),
            "level_ms": [round(x, 3) for x in st["level_ms"]],
            "tuning": ctx_tuning,
            "autotune": ({"trials": tuned["trials"], "seconds": tuned["seconds"]} if tuned else None),
            "timed_region_s": round(t_all, 3),
        },
        "single_root": single,
        "batch_vs_single_root": (round(gteps / single["gteps_hmean_wall"], 3) if single else None),
        "graph": {"generate_ms": round(info["generate_ms"], 2),
                  "construct_ms": round(info["construct_ms"], 2)},
        "serve": None,
    }
    ctx.batch_end()
    ctx.close()
    if not a.no_serve:
        out["serve"] = serve_leg(a.serve_scales, a.serve_rates, a.serve_max_queries, a.edgefactor, a.serve_policy)
    if not a.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(a.cpu_threads or baseline_threads(), a.cpu_scale)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
