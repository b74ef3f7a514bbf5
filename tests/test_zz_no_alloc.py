"""g5_validate_pred_noalloc (no device allocation, one scratch array per GPU)
against g5_validate_pred through the C ABI: identical validity, per-rule
violation counts and traversed-edge counts, for valid pred arrays and for
arrays with planted mistakes, on two calls in a row of which the second has
the smaller input, and G5_ERR_LIMIT beyond the scratch array's size."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _graph(scale):
    import graph500_amd
    c = graph500_amd.G5Context(0)
    c.generate_kronecker(scale, 16, 2, 3)
    c.construct(2, release_edges=True)
    c.prepare_bfs()
    return c


def _cases(c, root):
    """The run's own pred array and three planted mistakes."""
    c.run_bfs(root)
    pred = c.get_pred()
    out = [("valid", pred)]
    reached = np.flatnonzero((pred != -1) & (np.arange(len(pred)) != root))
    a = pred.copy()
    v = int(reached[len(reached) // 2])
    a[v] = (a[v] & ((1 << 48) - 1)) | ((((a[v] >> 48) & 0xFFFF) + 1) << 48)  # level one too deep
    out.append(("depth", a))
    b = pred.copy()
    b[int(reached[7])] = -1  # a reached vertex dropped: its edges join reached and unreached
    out.append(("dropped", b))
    d = pred.copy()
    d[int(reached[3])] = len(pred) + 5  # parent out of range
    out.append(("range", d))
    return out


def _compare(c, root):
    for name, pred in _cases(c, root):
        ref = c.validate_pred(root, pred)
        got = c.validate_pred_noalloc(root, pred)
        assert got == ref, (name, ref, got)
        if name == "valid":
            assert got[0] and got[2] == c.teps_edges()
        else:
            assert not got[0], name


def test_noalloc_validation_equals_allocating_entry_point():
    import graph500_amd
    big = _graph(16)
    small = _graph(12)
    try:
        for c, tag in ((big, "S16"), (small, "S12"), (big, "S16 again")):
            roots = [int(r) for r in c.find_roots(3)]
            say("noalloc validate", tag, roots)
            for r in roots:
                _compare(c, r)
    finally:
        big.close()
        small.close()
    # beyond the scratch array: a clear error, and the allocating form still works
    c = _graph(23)
    try:
        root = int(c.find_roots(1)[0])
        c.run_bfs(root)
        pred = c.get_pred()
        with pytest.raises(graph500_amd.G5Error) as ex:
            c.validate_pred_noalloc(root, pred)
        assert ex.value.code == -5
        ok, _, cnt = c.validate_pred(root, pred)
        assert ok and cnt == c.teps_edges()
    finally:
        c.close()
