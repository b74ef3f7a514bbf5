"""GPU validator (g5_validate, g5_validate_pred) against the oracle's
restatement of BfsValidation::validate (mpi/validate.hpp:506-802).

Correct runs must pass with zero violations and the traversed-edge count of
the run; pred arrays with a planted mistake (made on the host, in the input
array) must fail on the GPU exactly when they fail in the oracle, with the
violation counted under the rule the mistake breaks.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LEVEL_SHIFT = 48


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g16(oracle_mod):
    import graph500_amd
    say("validate: generate+construct SCALE 16")
    c = graph500_amd.G5Context(0)
    c.generate_kronecker(16, 16, 2, 3)
    edges = c.edges_download()
    c.construct(2)
    c.prepare_bfs()
    oc = oracle_mod.Csr(edges)
    yield c, edges, oc
    c.close()


def test_valid_runs_pass(g16, oracle_mod):
    c, edges, oc = g16
    for r in c.find_roots(16).tolist():
        c.run_bfs(r)
        ok, err = c.validate()
        assert ok and not any(err.values()), (r, err)
        pred = c.get_pred()
        ok2, err2, visited = c.validate_pred(r, pred)
        assert ok2 and not any(err2.values()), (r, err2)
        assert visited == c.teps_edges()


def _pred_of(x):
    return (x << 16) >> 16


def _plant(c, oc, pred, root, kind):
    """Returns a copy of pred with one mistake of the given kind."""
    p = pred.copy()
    lv = np.where(p == -1, -1, p >> LEVEL_SHIFT)
    a = oc.arrays()
    offs, cols, new2old, old2new = a["offsets"], a["cols"], a["new2old"], a["old2new"]
    n = len(p)

    def nbrs(v):
        nv = old2new[v]
        return set(new2old[cols[offs[nv]:offs[nv + 1]]].tolist())

    lvl2 = np.flatnonzero(lv == 2)
    lvl1 = np.flatnonzero(lv == 1)
    if kind == "parent_edge":
        # a level-2 vertex gets a level-1 parent that is not its neighbour
        for v in lvl2[:200]:
            nb = nbrs(int(v))
            cand = [int(u) for u in lvl1[:2000] if int(u) not in nb]
            if cand:
                p[v] = cand[0] | (2 << LEVEL_SHIFT)
                return p, "parent_edge"
        pytest.skip("no non-adjacent level-1 vertex found")
    if kind == "depth":
        v = int(lvl2[0])
        p[v] = _pred_of(p[v]) | (3 << LEVEL_SHIFT)
        return p, "depth"
    if kind == "unreached":
        v = int(lvl2[0])
        p[v] = -1
        return p, "edge"
    if kind == "range":
        v = int(lvl2[0])
        p[v] = (n + 5) | (2 << LEVEL_SHIFT)
        return p, "range"
    if kind == "root":
        p[root] = int(lvl1[0])
        return p, "root"
    if kind == "self_parent":
        v = int(lvl2[0])
        p[v] = v | (2 << LEVEL_SHIFT)
        return p, "self_parent"
    if kind == "no_parent_with_depth":
        v = int(lvl2[0])
        p[v] = ((1 << LEVEL_SHIFT) - 1) | (2 << LEVEL_SHIFT)
        return p, "depth"
    raise AssertionError(kind)


@pytest.mark.parametrize("kind", ["parent_edge", "depth", "unreached", "range", "root", "self_parent",
                                  "no_parent_with_depth"])
def test_planted_mistakes_fail_like_oracle(g16, oracle_mod, kind):
    c, edges, oc = g16
    r = int(c.find_roots(1)[0])
    c.run_bfs(r)
    pred = c.get_pred()
    bad, rule = _plant(c, oc, pred, r, kind)
    say("planted", kind, "root", r)
    ok_o, _, msg = oracle_mod.validate(edges, oc.max_used_vertex() + 1, r, bad)
    assert not ok_o, kind
    ok, err, _ = c.validate_pred(r, bad)
    assert not ok, (kind, err, msg)
    assert err[rule] >= 1, (kind, err, msg)


def test_validate_pred_shape_and_root_checks(g16):
    import graph500_amd
    c, _, _ = g16
    with pytest.raises(ValueError):
        c.validate_pred(0, np.zeros(5, np.int64))
    with pytest.raises(graph500_amd.G5Error):
        c.validate_pred(-1, np.full(c.pred_size(), -1, np.int64))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_small_graph_modes_validate(oracle_mod, mode):
    """Each reorder mode, roots with and without edges (mode 0 keeps rows for
    vertices without edges)."""
    import graph500_amd
    v0 = np.array([0, 1, 2, 2, 5, 5, 6, 7, 7, 9, 3, 3, 12])
    v1 = np.array([1, 2, 0, 2, 6, 6, 7, 5, 7, 9, 3, 4, 13])
    e = oracle_mod.pack_edges(v0, v1)
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(mode)
        c.prepare_bfs()
        for r in [0, 5, 3, 9, 8, 12]:
            c.run_bfs(r)
            ok, err = c.validate()
            assert ok, (mode, r, err)
            ok2, err2, visited = c.validate_pred(r, c.get_pred())
            assert ok2, (mode, r, err2)
            assert visited == c.teps_edges(), (mode, r)


def test_validate_s20_all_roots(oracle_mod):
    import graph500_amd
    say("validate: SCALE 20")
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(20, 16, 2, 3)
        c.construct(2)
        c.prepare_bfs()
        for r in c.find_roots(16).tolist():
            c.run_bfs(r)
            ok, err = c.validate()
            assert ok, (r, err)
