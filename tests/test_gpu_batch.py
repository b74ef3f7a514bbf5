"""GPU parity of the batched (64 roots per sweep) BFS: every root's level
array equals the single-root result (reference fixtures, oracle), traversed
edge counts are identical, and the batched byte model equals the oracle's."""
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


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def s16ctx():
    import graph500_amd
    c = graph500_amd.G5Context(0)
    c.generate_kronecker(16)
    c.construct(2)
    c.batch_prepare()
    yield c
    c.close()


# push while the union's edges < nnz / alpha: default; pull only (vertex-parallel);
# pull only (edge-parallel sweep); push only; each through the persistent
# graph-replayed kernel and through per-level launches
@pytest.mark.parametrize("persistent", [1, 0])
@pytest.mark.parametrize("alpha,sweep,filt,beta", [(16.0, 0.25, 0, 50.0), (16.0, 0.25, 1, 0.0), (1e12, 0.0, 0, 0.0),
                                                   (1e12, 0.0, 1, 0.0), (1e12, 0.0, 2, 0.0), (1e12, 2.0, 0, 0.0),
                                                   (1e-9, 0.25, 0, 0.0)])
def test_batch_64_roots_match_reference_s16(s16ctx, alpha, sweep, filt, beta, persistent):
    g = gold("ref_s16_64roots.json")
    roots = [p["root"] for p in g["per_root"]]
    s16ctx.batch_set_tuning(alpha=alpha, sweep_below=sweep, persistent=persistent, pull_filter=filt, beta=beta,
                            lane_rule=0)
    say("batch s16 alpha", alpha, "sweep", sweep, "filter", filt, "beta", beta, "persistent", persistent)
    s16ctx.run_bfs_batch(roots)
    st = s16ctx.batch_stats()
    if alpha > 1e6:
        assert set(st["direction"]) == ({2} if sweep > 1 else {0}), st["direction"]
    if alpha < 1:
        assert set(st["direction"]) == {1}, st["direction"]
    teps = s16ctx.batch_teps_edges()
    for i, pr in enumerate(g["per_root"]):
        assert sha(s16ctx.batch_levels(i)) == pr["levels_sha256"], (alpha, pr["root"])
        assert teps[i] == pr["nedge"]
    st = s16ctx.batch_stats()
    assert sum(st["pairs"]) + 64 == sum(sum(p["level_histogram"]) for p in g["per_root"])
    s16ctx.batch_set_tuning(alpha=16.0, sweep_below=0.25, persistent=0, pull_filter=0, beta=50.0)


@pytest.mark.parametrize("compact", [1.01, 0.3])
def test_batch_compact_frontier_matches_reference_s16(s16ctx, compact):
    """Pull levels reading the compact frontier copy (always, or when the
    frontier spans under 30 % of the rows)."""
    g = gold("ref_s16_64roots.json")
    roots = [p["root"] for p in g["per_root"]]
    s16ctx.batch_set_tuning(alpha=1e12 if compact > 1 else 16.0, persistent=0, compact_below=compact, beta=0.0,
                            lane_rule=0)
    say("batch s16 compact", compact)
    s16ctx.run_bfs_batch(roots)
    teps = s16ctx.batch_teps_edges()
    for i, pr in enumerate(g["per_root"]):
        assert sha(s16ctx.batch_levels(i)) == pr["levels_sha256"], (compact, pr["root"])
        assert teps[i] == pr["nedge"]
    s16ctx.batch_set_tuning(alpha=16.0, compact_below=0.0, beta=50.0)


def _per_root_run(s16ctx, alpha, filt, mixed_push=1, unroll=0):
    g = gold("ref_s16_64roots.json")
    roots = [p["root"] for p in g["per_root"]]
    s16ctx.batch_set_tuning(alpha=alpha, persistent=0, lane_rule=1, pull_filter=filt, beta=0.0, sweep_below=0.0,
                            mixed_push=mixed_push, pull_unroll=unroll)
    say("batch s16 per-root rule alpha", alpha, "filter", filt, "mixed_push", mixed_push, "unroll", unroll)
    try:
        s16ctx.run_bfs_batch(roots)
        st = s16ctx.batch_stats()
        if alpha < 1e-6:
            assert set(st["direction"]) == {1}, st["direction"]
        if alpha > 1e6:
            assert st["direction"][0] == 1 and set(st["direction"][1:]) <= {0, 3}, st["direction"]
        teps = s16ctx.batch_teps_edges()
        for i, pr in enumerate(g["per_root"]):
            assert sha(s16ctx.batch_levels(i)) == pr["levels_sha256"], (alpha, pr["root"])
            assert teps[i] == pr["nedge"]
        assert sum(st["pairs"]) + 64 == sum(sum(p["level_histogram"]) for p in g["per_root"])
        return st["direction"]
    finally:
        s16ctx.batch_set_tuning(alpha=16.0, sweep_below=0.25, persistent=0, pull_filter=0, beta=50.0, lane_rule=0,
                                mixed_push=1, pull_unroll=0)


@pytest.mark.parametrize("mixed_push", [0, 1])
@pytest.mark.parametrize("alpha,filt", [(4.0, 0), (16.0, 1), (1e-9, 0), (1e12, 0), (0.05, 0)])
def test_batch_per_root_directions_match_reference_s16(s16ctx, alpha, filt, mixed_push):
    """lane_rule 1: each root pulls or pushes by its own frontier; a level
    with both runs the push of some roots (over the CSR chunks holding their
    frontier, or over their frontier as a queue) and the pull of the others.
    1e-9 pushes every root at every level, 1e12 pulls every root after its
    first level."""
    _per_root_run(s16ctx, alpha, filt, mixed_push)


@pytest.mark.parametrize("mixed_push", [0, 1])
@pytest.mark.parametrize("unroll", [0, 1, 2, 3, 4, 5, 6, 7])
def test_batch_mixed_level_pull_unroll_s16(s16ctx, unroll, mixed_push):
    """Every probe unrolling of the merging pull of a mixed level
    (k_ms_pull_lanes<4,2>, <8,2>, <8,4>, <4,4>, <2,1>, <1,1>, <2,2>, <1,2>)
    under both push forms:
    alpha 4 gives the S16 batch mixed levels."""
    d = _per_root_run(s16ctx, 4.0, 0, mixed_push, unroll)
    assert 3 in d, d


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("unroll", [0, 1, 2, 3, 4, 5, 6, 7])
def test_batch_pull_unroll_matches_reference_s16(s16ctx, unroll, rule):
    """Every probe unrolling of the per-level pull (k_ms_pull<4,2>, <8,2>,
    <8,4>, <4,4>, <2,1>, <1,1>, <2,2>, <1,2>) gives the reference's levels: rule 0 with an alpha
    that never pushes runs every level after the first as that pull; rule 1
    runs it once every root pulls."""
    g = gold("ref_s16_64roots.json")
    roots = [p["root"] for p in g["per_root"]]
    s16ctx.batch_set_tuning(lane_rule=rule, alpha=1e12 if rule == 0 else 128.0, pull_unroll=unroll, persistent=0,
                            beta=50.0, sweep_below=0.0, pull_filter=0, compact_below=0.0)
    say("batch s16 pull_unroll", unroll, "rule", rule)
    try:
        s16ctx.run_bfs_batch(roots)
        st = s16ctx.batch_stats()
        assert 0 in st["direction"], st["direction"]
        teps = s16ctx.batch_teps_edges()
        for i, pr in enumerate(g["per_root"]):
            assert sha(s16ctx.batch_levels(i)) == pr["levels_sha256"], (unroll, rule, pr["root"])
            assert teps[i] == pr["nedge"]
        assert sum(st["pairs"]) + 64 == sum(sum(p["level_histogram"]) for p in g["per_root"])
    finally:
        s16ctx.batch_set_tuning(alpha=16.0, sweep_below=0.25, persistent=0, pull_filter=0, beta=50.0, lane_rule=0,
                                pull_unroll=0)


def test_batch_per_root_levels_mix_directions_s16(s16ctx):
    """Some level of the S16 batch runs pushes and pulls together at alpha 4."""
    d = _per_root_run(s16ctx, 4.0, 0)
    assert 3 in d, d


def test_batch_model_bytes_match_oracle(s16ctx, oracle_mod):
    O = oracle_mod
    oc = O.Csr(O.generate_edges(16))
    a = oc.arrays()
    roots = oc.find_roots(64)
    s16ctx.batch_set_tuning(lane_rule=0, alpha=16.0, persistent=0)  # the oracle's model: every level a pull
    s16ctx.run_bfs_batch(roots)
    push, pull = s16ctx.batch_level_lanes()
    assert not push.any()
    levels_new = np.stack([s16ctx.batch_levels(i)[a["new2old"]] for i in range(64)])
    tot_o, per_o = oc.msbfs_bytes(levels_new)
    per, tot = s16ctx.batch_model_bytes()
    assert tot == tot_o
    assert per == per_o[1:len(per) + 1].tolist()


@pytest.mark.parametrize("alpha", [4.0, 128.0])
def test_batch_model_bytes_per_root_rule_match_oracle(s16ctx, oracle_mod, alpha):
    """Byte model of a run with per-root directions (levels where some roots
    push and the others pull) against the oracle given the run's lane masks."""
    O = oracle_mod
    oc = O.Csr(O.generate_edges(16))
    a = oc.arrays()
    roots = oc.find_roots(64)
    s16ctx.batch_set_tuning(lane_rule=1, alpha=alpha, persistent=0, beta=50.0)
    try:
        s16ctx.run_bfs_batch(roots)
        st = s16ctx.batch_stats()
        push, pull = s16ctx.batch_level_lanes()
        if alpha == 4.0:
            assert 3 in st["direction"], st["direction"]
            assert any(push[l] and pull[l] for l in range(1, st["num_levels"]))
        levels_new = np.stack([s16ctx.batch_levels(i)[a["new2old"]] for i in range(64)])
        tot_o, per_o = oc.msbfs_bytes(levels_new, pull.tolist(), push.tolist())
        per, tot = s16ctx.batch_model_bytes()
        assert per == per_o[1:len(per) + 1].tolist()
        assert tot == tot_o
    finally:
        s16ctx.batch_set_tuning(lane_rule=0, alpha=16.0)


@pytest.mark.parametrize("mode,rule", [(2, 1), (2, 0), (1, 1), (0, 1)])
def test_batch_partial_and_odd_roots(oracle_mod, mode, rule):
    """Duplicate and isolated roots, partial batches, under each reorder mode
    (mode 0 keeps rows without edges; the per-root rule then falls back to
    one direction per level)."""
    import graph500_amd
    O = oracle_mod
    e = O.generate_edges(12)
    oc = O.Csr(e)
    iso = [v for v in range(oc.num_verts) if not oc.has_edge(v)][:2]
    roots = list(oc.find_roots(5)) + iso + [int(oc.find_roots(1)[0])]  # duplicate root, isolated roots
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(mode)
        c.batch_prepare()
        c.batch_set_tuning(lane_rule=rule, alpha=4.0 if rule else 16.0)
        for rs in (roots, roots[:1]):
            say("batch odd roots mode", mode, "rule", rule, rs)
            c.run_bfs_batch(rs)
            teps = c.batch_teps_edges()
            for i, r in enumerate(rs):
                if oc.has_edge(int(r)):
                    lv, _, tr = oc.bfs(int(r), want_pred=False)
                    assert teps[i] == tr["teps_edges"]
                else:
                    lv = np.full(oc.num_verts, 0xFF, np.uint8)
                    lv[int(r)] = 0
                assert np.array_equal(c.batch_levels(i), lv), (i, r)


@pytest.mark.parametrize("scale,sweep,rule,beta", [(20, 0.25, 0, 50.0), (20, 2.0, 0, 50.0), (20, 0.0, 1, 50.0),
                                                   (20, 0.0, 0, 0.0), (20, 0.0, 1, 0.0), (24, 0.0, 1, 50.0)])
def test_batch_matches_single_root_gpu(scale, sweep, rule, beta):
    """beta 0 lets small tail frontiers push again after the list-driven
    tail pulls (whose frontier words and nonzero bits the push then reads).
    At SCALE 24 every one of the 64 lanes is checked against the reference's
    64-root run."""
    import graph500_amd
    g = gold("ref_s24_64roots.json" if scale == 24 else "ref_s%d_16roots.json" % scale)
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(scale)
        c.construct(2, release_edges=True)
        roots = c.find_roots(64)
        c.batch_prepare()
        c.batch_set_tuning(sweep_below=sweep, lane_rule=rule, persistent=0, beta=beta)
        say("batch s%d sweep %g lane_rule %d beta %g" % (scale, sweep, rule, beta))
        c.run_bfs_batch(roots)
        teps = c.batch_teps_edges()
        assert roots.tolist()[:len(g["roots"])] == g["roots"]
        for i, pr in enumerate(g["per_root"]):
            assert sha(c.batch_levels(i)) == pr["levels_sha256"]
            assert teps[i] == pr["nedge"]
        c.prepare_bfs()
        for i in range(len(g["per_root"]), 64, 7):
            c.run_bfs(int(roots[i]))
            assert np.array_equal(c.get_levels(), c.batch_levels(i)), i
            assert c.teps_edges() == teps[i]


@pytest.mark.parametrize("unroll,alpha,mixed_push,grid", [(0, 256.0, 0, 8192), (4, 256.0, 0, 8192),
                                                           (4, 256.0, 1, 8192), (0, 32.0, 1, 8192),
                                                           (4, 48.0, 1, 8192), (4, 64.0, 1, 32768),
                                                           (4, 64.0, 1, 65536)])
@pytest.mark.parametrize("scale", [20, 22])
def test_batch_tuned_combination_matches_reference(scale, unroll, alpha, mixed_push, grid):
    """Combinations the autotuner picks at SCALE 28 (per-root rule, beta 200,
    filter_from 0.001; alpha 256 with the chunk push of mixed levels, alpha
    32-64 with the queue push; probe unrolling 4 or 0; pull grids up to
    65,536 workgroups, light levels capped): the fixture roots' level hashes
    and edge counts, and single-root runs of roots 16-63."""
    import graph500_amd
    g = gold("ref_s%d_16roots.json" % scale)
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(scale)
        c.construct(2, release_edges=True)
        roots = c.find_roots(64)
        c.batch_prepare()
        c.batch_set_tuning(lane_rule=1, alpha=alpha, beta=200.0, filter_from=0.001, pull_unroll=unroll,
                           pull_grid=grid, sweep_below=0.0, persistent=0, pull_filter=0, compact_below=0.0,
                           mixed_push=mixed_push)
        say("batch s%d tuned combination, unroll %d alpha %g mixed_push %d grid %d" % (scale, unroll, alpha,
                                                                                      mixed_push, grid))
        c.run_bfs_batch(roots)
        teps = c.batch_teps_edges()
        for i, pr in enumerate(g["per_root"]):
            assert sha(c.batch_levels(i)) == pr["levels_sha256"], (scale, unroll, pr["root"])
            assert teps[i] == pr["nedge"]
        c.prepare_bfs()
        for i in range(16, 64, 5):
            c.run_bfs(int(roots[i]))
            assert np.array_equal(c.get_levels(), c.batch_levels(i)), i
            assert c.teps_edges() == teps[i]


def test_batch_sweep_and_csr_push_small_graphs(oracle_mod):
    """The edge-parallel sweep and the persistent CSR push on reorder modes
    1 and 2, on graphs whose rows
    are shorter than a chunk, longer than a chunk, and of degree one."""
    import graph500_amd
    O = oracle_mod
    n = 5000
    star = (np.concatenate([np.zeros(n - 1, np.int64), np.arange(n, n + 60)]),
            np.concatenate([np.arange(1, n), np.arange(n + 1, n + 61)]))
    rng = np.random.default_rng(3)
    rnd = (rng.integers(0, 3000, 20000), rng.integers(0, 3000, 20000))
    for v0, v1 in (star, rnd):
        e = O.pack_edges(v0, v1)
        oc = O.Csr(e)
        roots = [r for r in range(0, 3000, 97) if oc.has_edge(r)][:20]
        for mode in (1, 2):
            with graph500_amd.G5Context(0) as c:
                c.edges_append(e)
                c.construct(mode)
                c.batch_prepare()
                for persistent, alpha in ((0, 1e12), (1, 1e12), (0, 1e-9), (1, 1e-9)):
                    c.batch_set_tuning(alpha=alpha, sweep_below=2.0, persistent=persistent, beta=0.0, lane_rule=0)
                    say("sweep small mode", mode, len(e), "persistent", persistent, "alpha", alpha)
                    c.run_bfs_batch(roots)
                    teps = c.batch_teps_edges()
                    for i, r in enumerate(roots):
                        lv, _, tr = oc.bfs(int(r), want_pred=False)
                        assert np.array_equal(c.batch_levels(i), lv), (mode, r, persistent)
                        assert teps[i] == tr["teps_edges"]


def test_batch_graph_replays_with_changing_roots(oracle_mod):
    """The persistent kernel's hipGraph is replayed for successive batches
    with different root sets (and eager single-root work in between); every
    replay must equal the oracle."""
    import graph500_amd
    O = oracle_mod
    e = O.generate_edges(14)
    oc = O.Csr(e)
    cand = oc.find_roots(64 + 40)
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(2)
        c.prepare_bfs()
        c.batch_prepare()
        c.batch_set_tuning(persistent=1)
        for k, rs in enumerate((cand[:64], cand[40:], cand[7:20], cand[:64])):
            say("replay", k, len(rs))
            c.run_bfs_batch(rs)
            teps = c.batch_teps_edges()
            for i, r in enumerate(rs):
                lv, _, tr = oc.bfs(int(r), want_pred=False)
                assert np.array_equal(c.batch_levels(i), lv), (k, i, r)
                assert teps[i] == tr["teps_edges"], (k, i)
            c.run_bfs(int(rs[0]))  # eager work between replays
            assert np.array_equal(c.get_levels(), c.batch_levels(0))


def test_batch_persistent_deeper_than_its_levels(oracle_mod):
    """A path of 200 vertices needs ~200 levels: the persistent launch covers
    persist_levels of them and the host loop the rest."""
    import graph500_amd
    O = oracle_mod
    n = 200
    e = O.pack_edges(np.arange(n - 1), np.arange(1, n))
    oc = O.Csr(e)
    roots = [0, 50, 199, 100]
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(2)
        c.batch_prepare()
        for pl in (4, 16):
            c.batch_set_tuning(persistent=1, persist_levels=pl)
            say("deep path persist_levels", pl)
            c.run_bfs_batch(roots)
            for i, r in enumerate(roots):
                lv, _, _ = oc.bfs(int(r), want_pred=False)
                assert np.array_equal(c.batch_levels(i), lv), (pl, r)
            assert c.batch_stats()["num_levels"] > pl


def test_autotune_keeps_results(oracle_mod, tmp_path):
    """The autotuner's pick (measured on the GPU) still gives the oracle's
    levels; a second call reuses the cached pick."""
    import graph500_amd
    from graph500_amd import autotune
    O = oracle_mod
    e = O.generate_edges(14)
    oc = O.Csr(e)
    roots = oc.find_roots(64)
    cache = str(tmp_path / "tuning.json")
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(2)
        c.batch_prepare()
        say("autotune s14")
        res = autotune.autotune(c, roots, reps=1, budget_s=20, cache=cache)
        assert not res["cached"] and res["trials"] > 5
        assert c.batch_tuning() == res["tuning"]
        c.run_bfs_batch(roots)
        for i in range(0, 64, 9):
            lv, _, _ = oc.bfs(int(roots[i]), want_pred=False)
            assert np.array_equal(c.batch_levels(i), lv), i
        again = autotune.autotune(c, roots, cache=cache)
        assert again["cached"] and again["tuning"] == res["tuning"]
