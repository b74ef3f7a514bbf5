"""GPU tests of the serving loop (g5_server_*): continuous batching over
resident graphs.  Every answer must equal the single-root BFS levels of its
root (bit-exact) with the same traversed-edge count, whether the query ran
alone or in a mix, whatever the arrival pattern, lane pressure or answer
memory pressure."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def graphs():
    import graph500_amd
    ctxs = []
    for scale in (12, 14, 16):
        say("serve graph S%d" % scale)
        c = graph500_amd.G5Context(0)
        c.generate_kronecker(scale, 16, 2, 3)
        c.construct(2, release_edges=True)
        c.prepare_bfs()
        ctxs.append(c)
    yield ctxs
    for c in ctxs:
        c.close()


def _roots_with_edges(ctx, k, rng):
    n = ctx.pred_size()
    out = []
    while len(out) < k:
        v = int(rng.integers(0, n))
        if ctx.has_edge(v):
            out.append(v)
    return out


def _single(ctx, root):
    ctx.run_bfs(root)
    return ctx.get_levels(), ctx.teps_edges()


def _queries(specs):
    import graph500_amd
    q = np.zeros(len(specs), graph500_amd.QUERY_DTYPE)
    for i, (g, r, t, keep) in enumerate(specs):
        q[i] = (g, keep, r, t)
    return q


@pytest.mark.parametrize("lane_rule", [0, 1, 2, 3])
def test_serve_answers_equal_single_root(graphs, lane_rule):
    import graph500_amd
    rng = np.random.default_rng(7)
    specs, t = [], 0.0
    for i in range(160):
        g = int(rng.integers(0, 3))
        t += float(rng.exponential(2e-4))
        specs.append((g, _roots_with_edges(graphs[g], 1, rng)[0], t, 1))
    with graph500_amd.G5Server(graphs, lane_rule=lane_rule) as srv:
        say("serve 160 queries")
        res, rep = srv.run(_queries(specs), timeout_s=120)
        assert rep["completed"] == 160
        assert (res["status"] == 2).all()
        assert (res["done_s"] >= res["admit_s"]).all()
        for i, (g, r, _, _) in enumerate(specs):
            lv, e = _single(graphs[g], r)
            assert np.array_equal(srv.levels(res["request"][i], graphs[g].pred_size()), lv), (i, g, r)
            assert res["edges"][i] == e, (i, g, r)
            srv.release(res["request"][i])


def test_alone_and_in_the_mix_are_identical(graphs):
    import graph500_amd
    rng = np.random.default_rng(11)
    picks = [(g, _roots_with_edges(graphs[g], 1, rng)[0]) for g in (0, 1, 2, 2, 1)]
    with graph500_amd.G5Server(graphs) as srv:
        alone = []
        for g, r in picks:
            res, _ = srv.run(_queries([(g, r, 0.0, 1)]))
            alone.append((srv.levels(res["request"][0], graphs[g].pred_size()), res["edges"][0]))
            srv.release(res["request"][0])
        # the same queries inside a burst of 300 others, arriving in the middle of it
        specs, t = [], 0.0
        where = {}
        for i in range(300):
            t += float(rng.exponential(1e-4))
            g = int(rng.integers(0, 3))
            specs.append((g, _roots_with_edges(graphs[g], 1, rng)[0], t, 0))
            if i in (40, 90, 150, 151, 220):
                k = len(where)
                where[len(specs)] = k
                specs.append((picks[k][0], picks[k][1], t, 1))
        say("serve mix of", len(specs))
        res, rep = srv.run(_queries(specs), timeout_s=120)
        assert (res["status"] >= 2).all()
        assert rep["max_active_lanes"] > 1
        for i, k in where.items():
            g, r = picks[k]
            lv = srv.levels(res["request"][i], graphs[g].pred_size())
            assert np.array_equal(lv, alone[k][0]), (k, g, r)
            assert res["edges"][i] == alone[k][1]
            srv.release(res["request"][i])


def test_cohort_admission_keeps_answers(graphs):
    """With cohort admission (a busy graph waits for 16 free lanes or 2 ms)
    answers stay exact."""
    import graph500_amd
    rng = np.random.default_rng(5)
    specs, t = [], 0.0
    for i in range(200):
        t += float(rng.exponential(5e-5))
        g = int(rng.integers(0, 3))
        specs.append((g, _roots_with_edges(graphs[g], 1, rng)[0], t, 1 if i % 10 == 0 else 0))
    with graph500_amd.G5Server(graphs, min_admit=16, max_wait_s=0.002) as srv:
        res, _ = srv.run(_queries(specs), timeout_s=120)
        assert (res["status"] >= 2).all()
        for i in range(0, 200, 10):
            g, r = specs[i][0], specs[i][1]
            lv, e = _single(graphs[g], r)
            assert np.array_equal(srv.levels(res["request"][i], graphs[g].pred_size()), lv), i
            assert res["edges"][i] == e
            srv.release(res["request"][i])


def test_lane_and_memory_pressure(graphs):
    """200 simultaneous queries on one graph (more than its 64 lanes) with only
    5 answer arrays: queries wait for lanes and memory, all complete."""
    import graph500_amd
    rng = np.random.default_rng(3)
    g = 1
    roots = _roots_with_edges(graphs[g], 200, rng)
    n = graphs[g].pred_size()
    isolated = next(v for v in range(n) if not graphs[g].has_edge(v))
    roots[17] = isolated          # a root without edges
    roots[18] = roots[19]         # a duplicate root
    with graph500_amd.G5Server(graphs, answer_buffers=5) as srv:
        keep = [1 if i in (0, 17, 18, 199) else 0 for i in range(200)]
        res, rep = srv.run(_queries([(g, r, 0.0, k) for r, k in zip(roots, keep)]), timeout_s=120)
        assert (res["status"] >= 2).all() and rep["max_active_lanes"] <= 5
        for i in (0, 17, 18, 199):
            if i == 17:
                lv = np.full(n, 0xFF, np.uint8)
                lv[isolated] = 0
                e = 0
            else:
                lv, e = _single(graphs[g], roots[i])
            assert np.array_equal(srv.levels(res["request"][i], n), lv), i
            assert res["edges"][i] == e
            srv.release(res["request"][i])


def test_serve_errors(graphs):
    import graph500_amd
    with graph500_amd.G5Server(graphs[:2]) as srv:
        with pytest.raises(graph500_amd.G5Error):
            srv.run(_queries([(2, 5, 0.0, 0)]))  # no such graph
        with pytest.raises(graph500_amd.G5Error):
            srv.run(_queries([(0, 5, 0.2, 0), (0, 6, 0.1, 0)]))  # arrivals out of order
        res, _ = srv.run(_queries([(0, _roots_with_edges(graphs[0], 1, np.random.default_rng(1))[0], 0.0, 0)]))
        with pytest.raises(graph500_amd.G5Error):
            srv.levels(res["request"][0], graphs[0].pred_size())  # not kept
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(10)
        c.construct(0)
        with pytest.raises(graph500_amd.G5Error):
            graph500_amd.G5Server([c])  # reorder mode 0 graphs are not served
