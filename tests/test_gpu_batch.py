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


@pytest.mark.parametrize("alpha", [16.0, 1e-9, 1e12])  # default, pull only, push only
def test_batch_64_roots_match_reference_s16(s16ctx, alpha):
    g = gold("ref_s16_64roots.json")
    roots = [p["root"] for p in g["per_root"]]
    s16ctx.batch_set_tuning(alpha=alpha)
    say("batch s16 alpha", alpha)
    s16ctx.run_bfs_batch(roots)
    teps = s16ctx.batch_teps_edges()
    for i, pr in enumerate(g["per_root"]):
        assert sha(s16ctx.batch_levels(i)) == pr["levels_sha256"], (alpha, pr["root"])
        assert teps[i] == pr["nedge"]
    st = s16ctx.batch_stats()
    assert sum(st["pairs"]) + 64 == sum(sum(p["level_histogram"]) for p in g["per_root"])
    s16ctx.batch_set_tuning(alpha=16.0)


def test_batch_model_bytes_match_oracle(s16ctx, oracle_mod):
    O = oracle_mod
    oc = O.Csr(O.generate_edges(16))
    a = oc.arrays()
    roots = oc.find_roots(64)
    s16ctx.run_bfs_batch(roots)
    levels_new = np.stack([s16ctx.batch_levels(i)[a["new2old"]] for i in range(64)])
    tot_o, per_o = oc.msbfs_bytes(levels_new)
    per, tot = s16ctx.batch_model_bytes()
    assert tot == tot_o
    assert per == per_o[1:len(per) + 1].tolist()


def test_batch_partial_and_odd_roots(oracle_mod):
    import graph500_amd
    O = oracle_mod
    e = O.generate_edges(12)
    oc = O.Csr(e)
    iso = [v for v in range(oc.num_verts) if not oc.has_edge(v)][:2]
    roots = list(oc.find_roots(5)) + iso + [int(oc.find_roots(1)[0])]  # duplicate root, isolated roots
    with graph500_amd.G5Context(0) as c:
        c.edges_append(e)
        c.construct(2)
        c.batch_prepare()
        for rs in (roots, roots[:1]):
            say("batch odd roots", rs)
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


@pytest.mark.parametrize("scale", [20])
def test_batch_matches_single_root_gpu(scale):
    import graph500_amd
    g = gold("ref_s%d_16roots.json" % scale)
    with graph500_amd.G5Context(0) as c:
        c.generate_kronecker(scale)
        c.construct(2, release_edges=True)
        roots = c.find_roots(64)
        c.batch_prepare()
        say("batch s%d" % scale)
        c.run_bfs_batch(roots)
        teps = c.batch_teps_edges()
        for i, pr in enumerate(g["per_root"]):
            assert sha(c.batch_levels(i)) == pr["levels_sha256"]
            assert teps[i] == pr["nedge"]
        c.prepare_bfs()
        for i in range(16, 64, 7):
            c.run_bfs(int(roots[i]))
            assert np.array_equal(c.get_levels(), c.batch_levels(i)), i
            assert c.teps_edges() == teps[i]
