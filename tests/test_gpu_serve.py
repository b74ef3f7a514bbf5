"""GPU tests of the serving loop (g5_server_*): continuous batching over
resident graphs.  Every answer must equal the oracle's BFS levels of its root
(bit-exact) with the oracle's traversed-edge count (mpi/validate.hpp:703-721),
whether the query ran alone or in a mix, whatever the arrival pattern, lane
pressure or answer memory pressure; on the SCALE 18 and 22 graphs the 16
roots of the reference run are served inside a mix and checked against the
reference's own level hashes and edge counts."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gold(name):
    return json.load(open(os.path.join(GOLD, name)))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def graphs(oracle_mod):
    """Three resident graphs (SCALE 12, 14, 16) and the oracle's CSR of each."""
    import graph500_amd
    ctxs, ocs = [], []
    for scale in (12, 14, 16):
        say("serve graph S%d" % scale)
        c = graph500_amd.G5Context(0)
        c.generate_kronecker(scale, 16, 2, 3)
        c.construct(2, release_edges=True)
        c.prepare_bfs()
        ctxs.append(c)
        ocs.append(oracle_mod.Csr(oracle_mod.generate_edges(scale)))
    yield ctxs, ocs
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


_memo = {}


def _oracle(oc, root):
    """(levels by original id, traversed edges) of the oracle's BFS; a root
    without edges is its own tree and traverses nothing."""
    key = (id(oc), root)
    if key not in _memo:
        if oc.has_edge(root):
            lv, _, tr = oc.bfs(root, want_pred=False)
            _memo[key] = (lv, tr["teps_edges"])
        else:
            lv = np.full(oc.num_verts, 0xFF, np.uint8)
            lv[root] = 0
            _memo[key] = (lv, 0)
    return _memo[key]


def _queries(specs):
    import graph500_amd
    q = np.zeros(len(specs), graph500_amd.QUERY_DTYPE)
    for i, (g, r, t, keep) in enumerate(specs):
        q[i] = (g, keep, r, t)
    return q


def _check_kept(srv, res, specs, ctxs, ocs, which):
    for i in which:
        g, r = specs[i][0], specs[i][1]
        lv, e = _oracle(ocs[g], r)
        assert np.array_equal(srv.levels(res["request"][i], ctxs[g].pred_size()), lv), (i, g, r)
        assert res["edges"][i] == e, (i, g, r)
        srv.release(res["request"][i])


@pytest.mark.parametrize("lane_rule", [0, 1, 2, 3])
def test_serve_answers_equal_oracle(graphs, lane_rule):
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(7)
    specs, t = [], 0.0
    for i in range(160):
        g = int(rng.integers(0, 3))
        t += float(rng.exponential(2e-4))
        specs.append((g, _roots_with_edges(ctxs[g], 1, rng)[0], t, 1))
    with graph500_amd.G5Server(ctxs, lane_rule=lane_rule) as srv:
        say("serve 160 queries, lane rule", lane_rule)
        res, rep = srv.run(_queries(specs), timeout_s=120)
        assert rep["completed"] == 160
        assert (res["status"] == 2).all()
        assert (res["done_s"] >= res["admit_s"]).all()
        _check_kept(srv, res, specs, ctxs, ocs, range(160))


def test_alone_and_in_the_mix_are_identical(graphs):
    """Five queries served alone, then again inside a burst of 300 others
    arriving around them: the same levels and edge counts, the oracle's."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(11)
    picks = [(g, _roots_with_edges(ctxs[g], 1, rng)[0]) for g in (0, 1, 2, 2, 1)]
    with graph500_amd.G5Server(ctxs) as srv:
        alone = []
        for g, r in picks:
            res, _ = srv.run(_queries([(g, r, 0.0, 1)]))
            alone.append((srv.levels(res["request"][0], ctxs[g].pred_size()), res["edges"][0]))
            srv.release(res["request"][0])
            lv, e = _oracle(ocs[g], r)
            assert np.array_equal(alone[-1][0], lv) and alone[-1][1] == e, (g, r)
        specs, t = [], 0.0
        where = {}
        for i in range(300):
            t += float(rng.exponential(1e-4))
            g = int(rng.integers(0, 3))
            specs.append((g, _roots_with_edges(ctxs[g], 1, rng)[0], t, 0))
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
            lv = srv.levels(res["request"][i], ctxs[g].pred_size())
            assert np.array_equal(lv, alone[k][0]), (k, g, r)
            assert res["edges"][i] == alone[k][1]
            srv.release(res["request"][i])


@pytest.mark.parametrize("min_admit,max_wait_s", [(1, 0.0), (4, 0.002), (8, 0.005), (16, 0.002)])
def test_queries_join_searches_in_flight(graphs, min_admit, max_wait_s):
    """Continuous batching: with a cohort threshold below 64, queries are
    admitted at level boundaries of graphs whose other searches are still
    running (admitted_busy > 0, in the report and in the per-graph counters),
    and every answer is the oracle's.  The load (50,000 queries/s, four in
    five on the SCALE 16 graph) lets a cohort of min_admit gather while a
    search of that graph is in flight."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(21 + min_admit)
    specs, t = [], 0.0
    for i in range(240):
        t += float(rng.exponential(2e-5))
        g = 2 if rng.random() < 0.8 else int(rng.integers(0, 2))
        specs.append((g, _roots_with_edges(ctxs[g], 1, rng)[0], t, 1 if i % 3 == 0 else 0))
    with graph500_amd.G5Server(ctxs, min_admit=min_admit, max_wait_s=max_wait_s) as srv:
        say("serve continuous admission min_admit %d max_wait %g" % (min_admit, max_wait_s))
        res, rep = srv.run(_queries(specs), timeout_s=120)
        assert (res["status"] >= 2).all()
        assert rep["admitted_busy"] > 0, rep
        assert rep["admit_events"] > 1
        busy = sum(srv.graph_stats(g)["admitted_busy"] for g in range(3))
        idle = sum(srv.graph_stats(g)["admitted_idle"] for g in range(3))
        assert busy == rep["admitted_busy"] and busy + idle == 240
        _check_kept(srv, res, specs, ctxs, ocs, range(0, 240, 3))


def test_cohort_admission_keeps_answers(graphs):
    """With cohort admission (a busy graph waits for 16 free lanes or 2 ms)
    answers stay exact."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(5)
    specs, t = [], 0.0
    for i in range(200):
        t += float(rng.exponential(5e-5))
        g = int(rng.integers(0, 3))
        specs.append((g, _roots_with_edges(ctxs[g], 1, rng)[0], t, 1 if i % 10 == 0 else 0))
    with graph500_amd.G5Server(ctxs, min_admit=16, max_wait_s=0.002) as srv:
        res, _ = srv.run(_queries(specs), timeout_s=120)
        assert (res["status"] >= 2).all()
        _check_kept(srv, res, specs, ctxs, ocs, range(0, 200, 10))


@pytest.mark.parametrize("solo", [1, 0])
def test_lone_queries_single_root_path(graphs, solo):
    """Queries spaced 15 ms apart find their graph idle: with solo 1 each is
    answered by the single-root kernels (solo count = queries), with solo 0
    each takes a lane; the answers are the oracle's either way."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(31)
    specs = []
    for i in range(12):
        g = i % 3
        specs.append((g, _roots_with_edges(ctxs[g], 1, rng)[0], 0.015 * i, 1 if i % 2 == 0 else 0))
    with graph500_amd.G5Server(ctxs, solo=solo) as srv:
        say("serve lone queries, solo", solo)
        res, rep = srv.run(_queries(specs), timeout_s=60)
        assert (res["status"] == 2).all()
        n_solo = sum(srv.graph_stats(g)["solo"] for g in range(3))
        assert n_solo == (12 if solo else 0), n_solo
        assert rep["admitted_busy"] == 0
        for i, (g, r, _, _) in enumerate(specs):
            assert res["edges"][i] == _oracle(ocs[g], r)[1], (i, g, r)
        _check_kept(srv, res, specs, ctxs, ocs, range(0, 12, 2))


@pytest.mark.parametrize("solo,slots", [(4, 1), (2, 1), (4, 3), (4, 2)])
def test_short_queue_single_root_path(graphs, solo, slots):
    """Three queries arrive at once on each of two idle graphs: with solo 4
    (a queue of three is short) all six are answered by the single-root
    kernels, one after another with one slot per graph or overlapping with
    two or three; with solo 2 each graph batches its three into lanes.  The
    answers are the oracle's either way."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(37)
    specs = [(g, r, 0.0, 1) for g in (1, 2) for r in _roots_with_edges(ctxs[g], 3, rng)]
    with graph500_amd.G5Server(ctxs, solo=solo, solo_slots=slots) as srv:
        say("serve short queues, solo", solo, "slots", slots)
        res, rep = srv.run(_queries(specs), timeout_s=60)
        assert (res["status"] == 2).all()
        n_solo = sum(srv.graph_stats(g)["solo"] for g in range(3))
        assert n_solo == (6 if solo == 4 else 0), n_solo
        # runs started while another of the same graph was going on: the
        # second and third of each graph with three slots, at least the second
        # with two, none with one
        over = [srv.graph_stats(g)["solo_overlap"] for g in (1, 2)]
        if solo == 4 and slots == 3:
            assert over == [2, 2], over
        elif solo == 4 and slots == 2:
            assert all(o >= 1 for o in over), over
        else:
            assert over == [0, 0], over
        for i, (g, r, _, _) in enumerate(specs):
            assert res["edges"][i] == _oracle(ocs[g], r)[1], (i, g, r)
        _check_kept(srv, res, specs, ctxs, ocs, range(len(specs)))


def test_run_after_a_timed_out_run(graphs):
    """A run that times out (possibly with a single-root run still going)
    leaves the server usable: the next run answers exactly its own queries."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(41)
    with graph500_amd.G5Server(ctxs) as srv:
        with pytest.raises(graph500_amd.G5Error):
            srv.run(_queries([(2, _roots_with_edges(ctxs[2], 1, rng)[0], 0.0, 0)]), timeout_s=1e-5)
        specs = [(0, _roots_with_edges(ctxs[0], 1, rng)[0], 0.0, 1), (1, _roots_with_edges(ctxs[1], 1, rng)[0], 0.0, 1)]
        res, rep = srv.run(_queries(specs), timeout_s=60)
        assert rep["completed"] == 2 and (res["status"] == 2).all()
        _check_kept(srv, res, specs, ctxs, ocs, range(2))


def test_lane_and_memory_pressure(graphs):
    """200 simultaneous queries on one graph (more than its 64 lanes) with only
    5 answer arrays: queries wait for lanes and memory, all complete."""
    import graph500_amd
    ctxs, ocs = graphs
    rng = np.random.default_rng(3)
    g = 1
    roots = _roots_with_edges(ctxs[g], 200, rng)
    n = ctxs[g].pred_size()
    isolated = next(v for v in range(n) if not ctxs[g].has_edge(v))
    roots[17] = isolated          # a root without edges
    roots[18] = roots[19]         # a duplicate root
    with graph500_amd.G5Server(ctxs, answer_buffers=5) as srv:
        keep = [1 if i in (0, 17, 18, 199) else 0 for i in range(200)]
        specs = [(g, r, 0.0, k) for r, k in zip(roots, keep)]
        res, rep = srv.run(_queries(specs), timeout_s=120)
        assert (res["status"] >= 2).all() and rep["max_active_lanes"] <= 5
        assert rep["admitted_busy"] > 0  # arrays freed by finished queries go to waiting ones
        _check_kept(srv, res, specs, ctxs, ocs, (0, 17, 18, 199))


def test_reference_roots_served_in_a_mix():
    """Graphs of SCALE 18 and 22 resident together; the 16 roots of the
    reference run on each (tests/golden/ref_s18_16roots.json,
    ref_s22_16roots.json) arrive inside a Poisson stream of 400 other queries.
    Each kept answer's level array hashes to the reference's level hash for
    that root, and its traversed-edge count is the reference's."""
    import graph500_amd
    gs = [gold("ref_s18_16roots.json"), gold("ref_s22_16roots.json")]
    ctxs = []
    try:
        for scale, g in zip((18, 22), gs):
            say("serve graph S%d" % scale)
            c = graph500_amd.G5Context(0)
            c.generate_kronecker(scale, 16, 2, 3)
            c.construct(2, release_edges=True)
            assert c.graph_info()["num_active"] == g["num_active"]
            ctxs.append(c)
        rng = np.random.default_rng(9)
        specs, t, where = [], 0.0, {}
        fixture = [(gi, pr) for gi, g in enumerate(gs) for pr in g["per_root"]]
        order = rng.permutation(len(fixture))
        slots = set(rng.choice(400, size=len(fixture), replace=False).tolist())
        k = 0
        for i in range(400):
            t += float(rng.exponential(4e-4))
            gi = int(rng.random() < 0.7)  # most queries on the SCALE 22 graph
            specs.append((gi, _roots_with_edges(ctxs[gi], 1, rng)[0], t, 0))
            if i in slots:
                gi, pr = fixture[order[k]]
                k += 1
                where[len(specs)] = pr
                specs.append((gi, pr["root"], t, 1))
        with graph500_amd.G5Server(ctxs) as srv:
            say("serve reference roots in a mix of", len(specs))
            res, rep = srv.run(_queries(specs), timeout_s=120)
            assert (res["status"] >= 2).all()
            for i, pr in where.items():
                gi = specs[i][0]
                lv = srv.levels(res["request"][i], ctxs[gi].pred_size())
                assert sha(lv) == pr["levels_sha256"], (gi, pr["root"])
                assert res["edges"][i] == pr["nedge"], (gi, pr["root"])
                srv.release(res["request"][i])
        say("serve reference roots: %d checked, %d admitted into busy graphs" % (len(where), rep["admitted_busy"]))
    finally:
        for c in ctxs:
            c.close()


PF_NEDGE = {18: 4194250, 22: 67108130, 25: 536865258, 27: 2147462776}  # mpi/utils.hpp:3664-3708


def test_serve_config5_graphs_resident():
    """The serving configuration's graphs (SCALE 18, 22, 25 and 27, edgefactor
    16) resident together; 400 Poisson queries at 1,000 per second, graph
    picked with probability proportional to 2^-((27 - SCALE) / 2), uniform
    roots with edges.  Six kept answers per graph equal the single-root level
    array of their root, whose parents pass every validator rule on the GPU;
    the traversed-edge count equals the single-root run's, which is
    pf_nedge[SCALE] for a root in the giant component."""
    import graph500_amd
    scales = (18, 22, 25, 27)
    ctxs = []
    try:
        for sc in scales:
            say("serve config graph S%d" % sc)
            c = graph500_amd.G5Context(0)
            c.generate_kronecker(sc, 16, 2, 3)
            c.construct(2, release_edges=True)
            c.prepare_bfs()
            ctxs.append(c)
        w = np.array([2.0 ** (-(27 - sc) / 2.0) for sc in scales])
        w /= w.sum()
        rng = np.random.default_rng(27)
        nq = 400
        graph = rng.choice(len(scales), size=nq, p=w)
        arrival = np.cumsum(rng.exponential(1e-3, size=nq))
        keep = set()
        for gi in range(len(scales)):
            idx = np.flatnonzero(graph == gi)
            keep |= set(idx[:: max(1, len(idx) // 6)][:6].tolist())
        specs = [(int(graph[i]), _roots_with_edges(ctxs[graph[i]], 1, rng)[0], float(arrival[i]),
                  1 if i in keep else 0) for i in range(nq)]
        with graph500_amd.G5Server(ctxs) as srv:
            say("serve config: %d queries, %d kept" % (nq, len(keep)))
            res, rep = srv.run(_queries(specs), timeout_s=120)
            assert (res["status"] >= 2).all()
            say("serve config: admitted into busy graphs %d, admit events %d" % (rep["admitted_busy"],
                                                                                 rep["admit_events"]))
            for i in sorted(keep):
                gi, r = specs[i][0], specs[i][1]
                c = ctxs[gi]
                lv = srv.levels(res["request"][i], c.pred_size())
                c.run_bfs(r)
                ok, errs = c.validate()
                assert ok, (scales[gi], r, errs)
                assert np.array_equal(lv, c.get_levels()), (scales[gi], r)
                e = c.teps_edges()
                assert res["edges"][i] == e, (scales[gi], r)
                if e > PF_NEDGE[scales[gi]] // 2:  # the giant component
                    assert e == PF_NEDGE[scales[gi]], (scales[gi], r)
                srv.release(res["request"][i])
    finally:
        for c in ctxs:
            c.close()


def test_serve_errors(graphs):
    import graph500_amd
    ctxs, _ = graphs
    with graph500_amd.G5Server(ctxs[:2]) as srv:
        with pytest.raises(graph500_amd.G5Error):
            srv.run(_queries([(2, 5, 0.0, 0)]))  # no such graph
        with pytest.raises(graph500_amd.G5Error):
            srv.run(_queries([(0, 5, 0.2, 0), (0, 6, 0.1, 0)]))  # arrivals out of order
        res, _ = srv.run(_queries([(0, _roots_with_edges(ctxs[0], 1, np.random.default_rng(1))[0], 0.0, 0)]))
        with pytest.raises(graph500_amd.G5Error):
            srv.levels(res["request"][0], ctxs[0].pred_size())  # not kept
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(10)
        c.construct(0)
        with pytest.raises(graph500_amd.G5Error):
            graph500_amd.G5Server([c])  # reorder mode 0 graphs are not served
    with pytest.raises(graph500_amd.G5Error):
        graph500_amd.G5Server(ctxs[:1], solo=65)  # solo is 0..64
    with pytest.raises(graph500_amd.G5Error):
        graph500_amd.G5Server(ctxs[:1], solo_slots=9)  # at most 8 slots


def test_served_context_refuses_rebuild_and_destroy():
    """While a server serves a context's graph, g5_construct and g5_destroy
    fail with G5_ERR_STATE and leave the graph in place; once the server is
    destroyed both work again."""
    import graph500_amd
    c = graph500_amd.G5Context(0)
    c.generate_kronecker(10)
    c.construct(2)
    srv = graph500_amd.G5Server([c])
    with pytest.raises(graph500_amd.G5Error) as ex:
        c.construct(2)
    assert ex.value.code == -2
    with pytest.raises(graph500_amd.G5Error) as ex:
        c.close()
    assert ex.value.code == -2
    root = int(c.find_roots(1)[0])
    res, _ = srv.run(_queries([(0, root, 0.0, 0)]))  # the server still works on the kept graph
    assert res["status"][0] == 2 and res["edges"][0] > 0
    srv.close()
    c.construct(2)
    c.close()
