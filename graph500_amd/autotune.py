"""Autotuner for the batched BFS (g5_batch_tuning).

Picks, for one resident graph and one set of roots, the launch shapes and
switch thresholds of g5_run_bfs_batch by measurement: coordinate descent over
each knob in turn (the others fixed at their best so far), every trial the
device time of a whole batch (minimum of `reps` runs).  Every configuration
computes the same levels (the parity tests run each level kind against the
oracle); the tuner also refuses a trial whose traversed-edge counts or whose
(root, vertex) pairs discovered at each level differ from the first run's:
the edge counts fix each root's reached set, the per-level pair counts (which
do not depend on the direction a level took) catch a level assigned wrongly.

Results are cached in a JSON file keyed by the graph (estimated SCALE,
active vertices, directed entries, reorder mode) and the library build
(its version string names the GPU architecture), so a second run on the same
graph reuses them.

    from graph500_amd import G5Context, autotune
    best = autotune.autotune(ctx, roots)          # sets the tuning on ctx
"""
import json
import os
import time

DEFAULT_CACHE = os.environ.get("G5_TUNING_CACHE",
                               os.path.join(os.path.expanduser("~"), ".cache", "graph500_amd", "tuning.json"))

# knob -> candidate values, in the order the descent visits them.  A knob
# "a,b" takes pairs: the direction rule and its alpha are tried together, as
# alpha means a different threshold under each rule.
SPACE = [
    ("pull_unroll", [0, 1, 2, 3, 4, 5, 6, 7]),
    ("mixed_push", [0, 1]),
    ("lane_rule,alpha", [(0, 1.0), (0, 2.0), (0, 4.0), (0, 8.0), (0, 16.0), (0, 64.0),
                         (1, 16.0), (1, 24.0), (1, 32.0), (1, 40.0), (1, 48.0), (1, 56.0), (1, 64.0), (1, 80.0),
                         (1, 96.0), (1, 128.0), (1, 192.0), (1, 256.0)]),
    ("persistent", [0, 1]),
    ("sweep_below", [0.0, 0.25]),
    ("beta", [0.0, 20.0, 50.0, 200.0]),
    ("pull_filter", [0, 1, 2]),
    ("filter_from", [0.001, 0.01, 0.05]),
    ("compact_below", [0.0, 0.25, 0.5]),
    ("pull_grid", [4096, 8192, 16384, 32768, 65536]),
    ("persist_waves", [4, 6, 8]),
]
# a second pass over the knobs that depend most on each other (the probe
# unrolling moves the best direction threshold and back)
REFINE = ["pull_unroll", "lane_rule,alpha"]


def graph_key(ctx):
    info = ctx.graph_info()
    return "%s|S%d|act%d|nnz%d|vr%d" % (ctx.version(), info["scale_estimate"], info["num_active"], info["nnz"],
                                        info["reorder_mode"])


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home is no reason to fail a run


def _measure(ctx, roots, reps):
    """Best device time of `reps` batches, and the last one's per-level pair counts."""
    best, pairs = float("inf"), None
    for _ in range(reps):
        ctx.run_bfs_batch(roots)
        st = ctx.batch_stats()
        best = min(best, st["device_ms"])
        pairs = list(st.get("pairs", []))
    return best, pairs


def autotune(ctx, roots, reps=2, budget_s=60.0, cache=DEFAULT_CACHE, use_cache=True, log=None, min_gain=0.003):
    """Tunes ctx's batched BFS for `roots` (batch_prepare must have run).
    A trial replaces the best so far only when it is faster by more than the
    fraction min_gain, so that run-to-run noise does not move the tuning.
    Returns {"tuning": {...}, "ms": best batch ms, "trials": n, "seconds": s,
    "cached": bool}.  Stops early when budget_s is spent."""
    key = graph_key(ctx)
    data = _load(cache) if use_cache else {}
    if use_cache and key in data:
        ctx.batch_set_tuning(**data[key]["tuning"])
        out = dict(data[key])
        out["cached"] = True
        return out
    t0 = time.time()
    ctx.run_bfs_batch(roots)  # warm-up (allocates per-level arrays)
    ref_edges = ctx.batch_teps_edges().tolist()
    cur = ctx.batch_tuning()
    best_ms, ref_pairs = _measure(ctx, roots, reps)
    trials = 1
    space = dict(SPACE)
    for knob, values in SPACE + [(k, space[k]) for k in REFINE]:
        if knob == "persist_waves" and not cur["persistent"]:
            continue
        names = knob.split(",")
        for v in values:
            setting = dict(zip(names, v if len(names) > 1 else (v,)))
            if all(cur[k] == x for k, x in setting.items()) or time.time() - t0 > budget_s:
                continue
            trial = dict(cur, **setting)
            ctx.batch_set_tuning(**trial)
            ms, pairs = _measure(ctx, roots, reps)
            trials += 1
            if ctx.batch_teps_edges().tolist() != ref_edges:
                raise RuntimeError("autotune: traversed-edge counts changed under %s" % trial)
            if pairs != ref_pairs:
                raise RuntimeError("autotune: pairs discovered per level changed under %s" % trial)
            if log:
                log("autotune %s=%s: %.3f ms (best %.3f)" % (knob, v, ms, best_ms))
            if ms < best_ms * (1.0 - min_gain):
                best_ms, cur = ms, trial
        ctx.batch_set_tuning(**cur)
    res = {"tuning": cur, "ms": best_ms, "trials": trials, "seconds": round(time.time() - t0, 2),
           "cached": False}
    if use_cache:
        data = _load(cache)
        data[key] = {k: res[k] for k in ("tuning", "ms", "trials", "seconds")}
        _save(cache, data)
    return res
