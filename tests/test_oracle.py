"""Pins the CPU oracle against known answers (CPU only).

Known answers used, per oracle function:
  make_mrg_seed           closed form of generator/graph_generator.hpp:327-333
  transition table        entries of the reference's generated table (golden)
  mrg_skip                composition and brute-force stepping (math)
  generate_edges          reference tuple-stream SHA-256 (golden), pf_nedge
  Csr                     reference counts (golden) + brute force on small graphs
  find_roots              reference roots (golden)
  bfs                     reference level-array SHA-256, directions, frontier
                          sizes (golden) + brute-force BFS on small graphs
  edge_weights            the compiled branch's 0xBEEF constant; the per-edge
                          draw against the MRG recurrence raised to the skip
                          exponent with Python integers (math)
  validate                pf_nedge (mpi/utils.hpp:3681) + one planted error per rule
  statistics              direct formulas of mpi/logfile.h:19-44
"""
import collections
import hashlib
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M = 2**31 - 1


def gold(name):
    return json.load(open(os.path.join(GOLD, name)))


@pytest.fixture(scope="module")
def s16(oracle_mod):
    O = oracle_mod
    e = O.generate_edges(16)
    return e, O.Csr(e)


def test_seed_expansion(oracle_mod):
    assert list(oracle_mod.make_mrg_seed(2, 3)) == [3, 1, 4, 1, 1]
    s1, s2 = 0x123456789ABCDEF0, 0x0FEDCBA987654321
    z = oracle_mod.make_mrg_seed(s1, s2)
    assert list(z) == [(s1 & 0x3FFFFFFF) + 1, ((s1 >> 30) & 0x3FFFFFFF) + 1, (s2 & 0x3FFFFFFF) + 1,
                       ((s2 >> 30) & 0x3FFFFFFF) + 1, ((s2 >> 60) << 4) + (s1 >> 60) + 1]


def test_transition_table_matches_reference_entries(oracle_mod):
    g = gold("mrg_table.json")
    for v in g["vectors"]:
        got = oracle_mod.transition_entry(v["byte"], v["value"]).tolist()
        assert got == v["entry"], (v["byte"], v["value"])


def _py_step(z):
    nz = (107374182 * z[0] + 104480 * z[4]) % M
    return [nz, z[0], z[1], z[2], z[3]]


def test_mrg_step_recurrence(oracle_mod):
    z = [3, 1, 4, 1, 1]
    cur = np.array(z, np.uint32)
    for _ in range(50):
        v, cur = oracle_mod.mrg_get_uint(cur)
        z = _py_step(z)
        assert v == z[0] and cur.tolist() == z


def test_mrg_skip_brute_force_and_composition(oracle_mod):
    O = oracle_mod
    z0 = np.array([3, 1, 4, 1, 1], np.uint32)
    # brute force: skip(k) == k plain steps
    zz = [int(x) for x in z0]
    for k in range(1, 700):
        zz = _py_step(zz)
        if k in (1, 2, 5, 255, 256, 257, 300, 511, 699):
            assert O.mrg_skip(z0, 0, 0, k).tolist() == zz, k
    # composition across all three 64-bit exponent words
    a = (0x1, 0xDEADBEEF12345678, 0x0123456789ABCDEF)
    b = (0x7, 0x00000000FFFFFFFF, 0xFEDCBA9876543210)
    lo = a[2] + b[2]
    mid = a[1] + b[1] + (lo >> 64)
    hi = a[0] + b[0] + (mid >> 64)
    both = O.mrg_skip(O.mrg_skip(z0, *a), *b)
    assert both.tolist() == O.mrg_skip(z0, hi, mid & (2**64 - 1), lo & (2**64 - 1)).tolist()
    # the period of the generator divides m^5 - 1: A^(m^5-1) = I
    e = M**5 - 1
    assert O.mrg_skip(z0, e >> 128, (e >> 64) & (2**64 - 1), e & (2**64 - 1)).tolist() == z0.tolist()


def test_random_doubles_in_unit_interval(oracle_mod):
    d = oracle_mod.make_random_numbers(1000, 2, 3, 0)
    assert ((d >= 0) & (d < 1)).all()
    # position p starts where the stream skipped by 2p draws starts
    d2 = oracle_mod.make_random_numbers(10, 2, 3, 37)
    assert np.array_equal(d[37:47], d2)


def test_scramble_is_a_permutation(oracle_mod):
    O = oracle_mod
    v0, v1 = O.scramble_values(2, 3)
    for scale in (4, 10):
        img = [O.lib().or_scramble(i, scale, int(v0), int(v1)) for i in range(1 << scale)]
        assert sorted(img) == list(range(1 << scale))


def test_generator_matches_reference_stream_s16(oracle_mod, s16):
    e, _ = s16
    g = gold("ref_s16_16roots.json")
    t = e.view(np.uint32).reshape(-1, 3)
    assert len(t) == g["num_tuples"] == 16 << 16
    assert t[:8].tolist() == g["tuples_head"] and t[-8:].tolist() == g["tuples_tail"]
    assert hashlib.sha256(t.tobytes()).hexdigest() == g["tuples_sha256"]


def test_generator_range_is_position_independent(oracle_mod):
    O = oracle_mod
    full = O.generate_edges(10)
    part = O.generate_edges(10, start=1000, end=3000)
    assert np.array_equal(full[1000:3000], part)


def test_csr_counts_match_reference_s16(s16):
    _, c = s16
    g = gold("ref_s16_16roots.json")
    assert c.scale_estimate == g["scale_estimate"] == 16
    assert c.num_active == g["num_active"]
    assert c.nnz == g["nnz"] == 1819202
    assert c.directed_before == 2096170 and c.num_self_loops == 491


def _brute_csr(v0, v1):
    n = 1 << (int(np.bitwise_or.reduce(np.concatenate([v0, v1]))).bit_length())
    deg = collections.Counter()
    for a, b in zip(v0, v1):
        if a != b:
            deg[a] += 1
            deg[b] += 1
    order = sorted(range(n), key=lambda v: (-deg[v], v))
    new = {o: i for i, o in enumerate(order)}
    adj = collections.defaultdict(set)
    for a, b in zip(v0, v1):
        if a != b:
            adj[new[a]].add(new[b])
            adj[new[b]].add(new[a])
    na = sum(1 for v in range(n) if deg[v] > 0)
    offs, cols = [0], []
    for v in range(na):
        cols += sorted(adj[v])
        offs.append(len(cols))
    return order, na, offs, cols


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_csr_brute_force_small(oracle_mod, seed):
    O = oracle_mod
    rng = np.random.default_rng(seed)
    m = 300
    v0 = rng.integers(0, 200, m)
    v1 = rng.integers(0, 200, m)
    v1[:20] = v0[:20]  # self loops
    v0[20:40] = v0[40:60]
    v1[20:40] = v1[40:60]  # duplicates
    c = O.Csr(O.pack_edges(v0, v1))
    order, na, offs, cols = _brute_csr(v0, v1)
    a = c.arrays()
    assert c.num_active == na
    assert a["new2old"].tolist() == order
    assert a["offsets"].tolist() == offs
    assert a["cols"].tolist() == cols


def test_find_roots_match_reference(s16):
    _, c = s16
    assert c.find_roots(16).tolist() == gold("ref_s16_16roots.json")["roots"]
    assert c.find_roots(64).tolist() == gold("ref_s16_64roots.json")["roots"]


def test_bfs_matches_reference_levels_schedule_s16(oracle_mod, s16):
    e, c = s16
    g = gold("ref_s16_64roots.json")
    for pr in g["per_root"]:
        lv, pred, tr = c.bfs(pr["root"])
        assert hashlib.sha256(lv.tobytes()).hexdigest() == pr["levels_sha256"], pr["root"]
        assert "".join("T" if d else "B" for d in tr["direction"]) == pr["directions"]
        assert tr["nq"] == pr["nq"]
        assert tr["teps_edges"] == pr["nedge"] == 1048570  # pf_nedge[16]


def _brute_bfs(n, v0, v1, root):
    adj = collections.defaultdict(set)
    for a, b in zip(v0, v1):
        adj[a].add(b)
        adj[b].add(a)
    lv = {root: 0}
    q = collections.deque([root])
    while q:
        u = q.popleft()
        for w in adj[u]:
            if w not in lv:
                lv[w] = lv[u] + 1
                q.append(w)
    out = np.full(n, 0xFF, np.uint8)
    for k, v in lv.items():
        out[k] = v
    return out


@pytest.mark.parametrize("scale", [6, 8, 10])
def test_bfs_brute_force_small_kronecker(oracle_mod, scale):
    O = oracle_mod
    e = O.generate_edges(scale)
    c = O.Csr(e)
    v0, v1 = O.edge_endpoints(e)
    for r in c.find_roots(8):
        lv, pred, tr = c.bfs(int(r))
        assert np.array_equal(lv, _brute_bfs(c.num_verts, v0, v1, int(r)))
        ok, cnt, msg = O.validate(e, c.max_used_vertex() + 1, int(r), pred)
        assert ok, msg
        assert cnt == tr["teps_edges"]


def test_bfs_path_graph_switches_and_long_diameter(oracle_mod):
    O = oracle_mod
    n = 200
    v0 = np.arange(n - 1)
    v1 = np.arange(1, n)
    c = O.Csr(O.pack_edges(v0, v1))
    lv, pred, tr = c.bfs(0)
    assert lv[: n].tolist() == list(range(n))


def test_validator_accepts_and_each_rule_rejects(oracle_mod, s16):
    O = oracle_mod
    e, c = s16
    root = int(c.find_roots(1)[0])
    lv, pred, tr = c.bfs(root)
    ng = c.max_used_vertex() + 1
    ok, cnt, msg = O.validate(e, ng, root, pred)
    assert ok and cnt == 1048570
    reached = np.nonzero(lv != 0xFF)[0]
    unreached = np.nonzero((lv == 0xFF) & (np.arange(len(lv)) < ng))[0]
    deep = [v for v in reached if lv[v] >= 2][0]
    cases = {}
    p = pred.copy(); p[deep] = ng + 5 | (int(lv[deep]) << 48); cases["range"] = p
    p = pred.copy(); p[root] = int(deep); cases["root"] = p
    p = pred.copy(); p[deep] = int(deep) | (int(lv[deep]) << 48); cases["self"] = p
    p = pred.copy(); p[deep] = (pred[deep] & ((1 << 48) - 1)) | ((int(lv[deep]) + 1) << 48); cases["depth"] = p
    p = pred.copy(); p[unreached[0]] = root | (1 << 48); cases["unreached_gets_parent"] = p
    p = pred.copy(); p[deep] = -1; cases["reached_marked_unreached"] = p
    # parent with correct depth but no edge: pick a vertex at the parent's level that is not adjacent
    v0, v1 = O.edge_endpoints(e)
    nbrs = set(v1[v0 == deep].tolist()) | set(v0[v1 == deep].tolist())
    cand = [v for v in np.nonzero(lv == lv[deep] - 1)[0] if v not in nbrs]
    p = pred.copy(); p[deep] = int(cand[0]) | (int(lv[deep]) << 48); cases["no_edge"] = p
    for name, p in cases.items():
        ok, _, msg = O.validate(e, ng, root, p)
        assert not ok, name
        assert msg, name


def test_statistics_formulas(oracle_mod):
    x = np.array([3.0, 1.0, 2.0, 9.0, 4.0])
    s = oracle_mod.statistics(x)
    assert s["min"] == 1 and s["max"] == 9 and s["median"] == 3
    assert s["mean"] == pytest.approx(x.mean())
    assert s["std"] == pytest.approx(x.std(ddof=1))
    assert s["q1"] == 2 and s["q3"] == 4


@pytest.mark.parametrize("scale", [8, 10])
def test_single_root_bytes_brute_force(oracle_mod, scale):
    """The single-root byte model of the oracle (the figure the GPU's
    g5_bfs_model_bytes is tested equal to) against a restatement from the
    definition, given the run's levels and schedule: a top-down level costs
    sum over its frontier of (4 + 8 + 4 deg) plus 4 per discovered vertex; a
    bottom-up level 3 n / 8 plus, per vertex with edges not visited before
    it, 8 + 4 k with k the neighbours probed in ascending order up to and
    including the first one visited before the level (the whole row if
    none); every level adds 1 byte per discovered vertex, and the run 1 byte
    for the root."""
    O = oracle_mod
    c = O.Csr(O.generate_edges(scale))
    a = c.arrays()
    off, cols, new2old = a["offsets"], a["cols"], a["new2old"]
    n, na = c.num_verts, c.num_active
    for r in c.find_roots(4):
        lv_orig, _, tr = c.bfs(int(r), want_pred=False)
        lv = lv_orig[new2old].astype(np.int64)  # by new id
        lv[lv == 0xFF] = 1 << 20
        total = 1.0
        for L, d in enumerate(tr["direction"], start=1):
            found = int((lv == L).sum())
            if d:  # top-down
                front = np.flatnonzero(lv[:na] == L - 1)
                b = sum(12.0 + 4.0 * (off[u + 1] - off[u]) for u in front) + 4.0 * found
            else:
                b = 3.0 * n / 8.0
                for v in range(na):
                    if lv[v] < L:
                        continue
                    row = cols[off[v]:off[v + 1]]
                    hit = np.flatnonzero(lv[row] < L)
                    k = int(hit[0]) + 1 if len(hit) else len(row)
                    b += 8.0 + 4.0 * k
            b += found
            assert tr["bytes"][L - 1] == pytest.approx(b), (int(r), L, d)
            total += b
        assert tr["total_bytes"] == pytest.approx(total)
        # a bottom-up level probes each found vertex's row up to a visited neighbour only
        assert any(d == 0 for d in tr["direction"]) or scale < 10


@pytest.mark.parametrize("masked", [False, True])
def test_msbfs_bytes_brute_force(oracle_mod, masked):
    """The batched byte model against a set-based restatement; masked: with
    per-level pulled / pushed lane sets (random, seeded), as a run with
    per-root directions reports them."""
    O = oracle_mod
    e = O.generate_edges(8)
    c = O.Csr(e)
    a = c.arrays()
    roots = c.find_roots(5)
    n = c.num_verts
    levels_new = np.full((len(roots), n), 0xFF, np.uint8)
    for i, r in enumerate(roots):
        lv, _, _ = c.bfs(int(r), want_pred=False)
        levels_new[i] = lv[a["new2old"]]
    L = int(levels_new[levels_new != 0xFF].max()) + 1
    pull = push = None
    if masked:
        rng = np.random.default_rng(4)
        pull = [int(x) for x in rng.integers(0, 1 << len(roots), L + 1)]
        push = [int(x) for x in rng.integers(0, 1 << len(roots), L + 1)]
    tot, per = c.msbfs_bytes(levels_new, pull, push)
    # brute force in python
    na = c.num_active
    exp = 0.0  # no per-(root, vertex) output byte: the frontier words are the output
    for l in range(1, L + 1):
        b = 24.0 * n
        for v in range(na):
            deg = int(a["offsets"][v + 1] - a["offsets"][v])
            if masked and any((push[l] >> r) & 1 and levels_new[r, v] == l - 1 for r in range(len(roots))):
                b += 8 + 12.0 * deg
            need = [r for r in range(len(roots)) if (levels_new[r, v] == 0xFF or levels_new[r, v] >= l)
                    and (not masked or (pull[l] >> r) & 1)]
            if not need:
                continue
            got, k = set(), 0
            for j in range(a["offsets"][v], a["offsets"][v + 1]):
                k += 1
                w = a["cols"][j]
                got |= {r for r in range(len(roots)) if levels_new[r, w] == l - 1}
                if set(need) <= got:
                    break
            b += 8 + 12.0 * k
        exp += b
    assert tot == pytest.approx(exp)


@pytest.mark.parametrize("scale", [14, 16])
def test_spec2012_graph_matches_reference(oracle_mod, scale):
    """generate_graph_spec2012 (initiator 5500/100, binary-tree initial
    tuples first, mpi/benchmark_helper.hpp:316-322): the tuple stream, the
    CSR counts, the roots and every root's levels, schedule and traversed
    edges equal the reference run's."""
    O = oracle_mod
    g = gold("ref_spec2012_s%d_16roots.json" % scale)
    e = O.generate_graph_spec2012(scale)
    t = e.view(np.uint32).reshape(-1, 3)
    assert len(t) == g["num_tuples"] == (16 << scale) + (1 << scale) - 1
    assert hashlib.sha256(t.tobytes()).hexdigest() == g["tuples_sha256"]
    c = O.Csr(e)
    assert c.num_active == g["num_active"] == 1 << scale  # the tree reaches every vertex
    assert c.nnz == g["nnz"]
    assert c.find_roots(16).tolist() == g["roots"]
    for pr in g["per_root"][:6]:
        lv, _, tr = c.bfs(pr["root"], want_pred=False)
        # the reference's pred arrays are rounded up to 2^16 entries (unreached)
        lv = np.concatenate([lv, np.full(g["pred_size"] - len(lv), 0xFF, np.uint8)])
        assert hashlib.sha256(lv.tobytes()).hexdigest() == pr["levels_sha256"], pr["root"]
        assert "".join("T" if d else "B" for d in tr["direction"]) == pr["directions"]
        assert tr["teps_edges"] == pr["nedge"] == len(t)


def _mrg_matrix_power(e):
    """A^e for the MRG step z1' = (107374182 z1 + 104480 z5) mod M, z_k' = z_(k-1)
    (generator/splittable_mrg.c:170-177), by square-and-multiply on Python ints."""
    A = [[107374182, 0, 0, 0, 104480], [1, 0, 0, 0, 0], [0, 1, 0, 0, 0], [0, 0, 1, 0, 0], [0, 0, 0, 1, 0]]

    def mul(X, Y):
        return [[sum(X[i][k] * Y[k][j] for k in range(5)) % M for j in range(5)] for i in range(5)]
    R = [[int(i == j) for j in range(5)] for i in range(5)]
    while e:
        if e & 1:
            R = mul(R, A)
        A = mul(A, A)
        e >>= 1
    return R


def test_edge_weights(oracle_mod):
    """generateWeight (generator/graph_generator.hpp:480-506): the compiled
    branch stores 0xBEEF; the per-edge form is one draw from the seed state
    advanced by 30 * 2^128 + 46 * 2^64 + edge_index steps, mod max_weight, + 1.
    A dropped skip word, a wrong word order or an off-by-one draw fails the
    brute-force comparison."""
    O = oracle_mod
    assert (O.edge_weights(5, 1000, 0) == 0xBEEF).all()
    z0 = [int(x) for x in O.make_mrg_seed(2, 3)]
    for idx in [0, 1, 255, 256, 65537, (1 << 40) + 7]:
        P = _mrg_matrix_power(30 * 2**128 + 46 * 2**64 + idx)
        z = [sum(P[i][k] * z0[k] for k in range(5)) % M for i in range(5)]
        draw = (107374182 * z[0] + 104480 * z[4]) % M
        for mw in [1, 255, 1000]:
            assert int(O.edge_weights(idx, 1, mw)[0]) == draw % mw + 1, (idx, mw)
    w = O.edge_weights(100, 5000, 255)
    assert w.min() >= 1 and w.max() <= 255 and len(np.unique(w)) > 200
