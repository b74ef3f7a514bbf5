"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library never
loads it.  Each wrapper names the reference function it restates (see
g5_oracle.h for file:line citations).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

EDGE_DTYPE = np.dtype([("v0_low", "<u4"), ("v1_low", "<u4"), ("high", "<u4")])
OR_MAX_LEVELS = 256


class OrCsr(ctypes.Structure):
    _fields_ = [
        ("scale_estimate", ctypes.c_int),
        ("num_verts", ctypes.c_int64),
        ("num_active", ctypes.c_int64),
        ("num_tuples", ctypes.c_int64),
        ("num_self_loops", ctypes.c_int64),
        ("directed_before", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("offsets", ctypes.POINTER(ctypes.c_uint64)),
        ("cols", ctypes.POINTER(ctypes.c_uint32)),
        ("new2old", ctypes.POINTER(ctypes.c_uint32)),
        ("old2new", ctypes.POINTER(ctypes.c_uint32)),
        ("degree_raw", ctypes.POINTER(ctypes.c_uint32)),
        ("self_loops", ctypes.POINTER(ctypes.c_uint32)),
    ]


class OrBfsTrace(ctypes.Structure):
    _fields_ = [
        ("num_levels", ctypes.c_int),
        ("direction", ctypes.c_int * OR_MAX_LEVELS),
        ("nq", ctypes.c_int64 * OR_MAX_LEVELS),
        ("bytes", ctypes.c_double * OR_MAX_LEVELS),
        ("edges_examined", ctypes.c_int64 * OR_MAX_LEVELS),
        ("total_bytes", ctypes.c_double),
        ("reached", ctypes.c_int64),
        ("teps_edges", ctypes.c_int64),
    ]


_lib = None


def build():
    """Compile liboracle.so (plain gcc, see oracle/Makefile)."""
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.or_make_mrg_seed.argtypes = [ctypes.c_uint64, ctypes.c_uint64, u32p]
    L.or_mrg_transition_entry.argtypes = [ctypes.c_int, ctypes.c_int, u32p]
    L.or_mrg_skip.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
    L.or_mrg_get_uint.argtypes = [u32p]
    L.or_mrg_get_uint.restype = ctypes.c_uint32
    L.or_mrg_get_double.argtypes = [u32p]
    L.or_mrg_get_double.restype = ctypes.c_double
    L.or_make_random_numbers.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    L.or_scramble_values.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint64)]
    L.or_scramble.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
    L.or_scramble.restype = ctypes.c_int64
    L.or_generate_edges.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_void_p]
    L.or_csr_build.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.or_csr_build.restype = ctypes.POINTER(OrCsr)
    L.or_csr_free.argtypes = [ctypes.POINTER(OrCsr)]
    L.or_csr_has_edge.argtypes = [ctypes.POINTER(OrCsr), ctypes.c_int64]
    L.or_find_max_used_vertex.argtypes = [ctypes.POINTER(OrCsr)]
    L.or_find_max_used_vertex.restype = ctypes.c_int64
    L.or_find_roots.argtypes = [ctypes.POINTER(OrCsr), ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.or_bfs.argtypes = [ctypes.POINTER(OrCsr), ctypes.c_int64, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.POINTER(OrBfsTrace)]
    L.or_initial_tree_edges.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                        ctypes.c_int64, ctypes.c_void_p]
    L.or_edge_weights.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_void_p]
    L.or_msbfs_bytes.argtypes = [ctypes.POINTER(OrCsr), ctypes.c_int, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
    L.or_msbfs_bytes.restype = ctypes.c_double
    L.or_validate.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                              ctypes.c_int64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                              ctypes.c_char_p, ctypes.c_int]
    L.or_statistics.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def make_mrg_seed(s1, s2):
    z = np.zeros(5, np.uint32)
    lib().or_make_mrg_seed(s1, s2, _u32p(z))
    return z


def transition_entry(byte_index, value):
    out = np.zeros(9, np.uint32)
    lib().or_mrg_transition_entry(byte_index, value, _u32p(out))
    return out


def mrg_skip(z, hi, mid, lo):
    z = np.ascontiguousarray(z, np.uint32).copy()
    lib().or_mrg_skip(_u32p(z), hi, mid, lo)
    return z


def mrg_get_uint(z):
    """Returns (value, new_state)."""
    z = np.ascontiguousarray(z, np.uint32).copy()
    v = lib().or_mrg_get_uint(_u32p(z))
    return v, z


def make_random_numbers(n, s1, s2, pos):
    out = np.zeros(n, np.float64)
    lib().or_make_random_numbers(n, s1, s2, pos, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def scramble_values(s1, s2):
    out = np.zeros(2, np.uint64)
    lib().or_scramble_values(s1, s2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def generate_edges(scale, edgefactor=16, s1=2, s2=3, a_num=5700, bc_num=1900, start=0, end=None):
    if end is None:
        end = edgefactor << scale
    out = np.zeros(end - start, EDGE_DTYPE)
    lib().or_generate_edges(scale, edgefactor, s1, s2, a_num, bc_num, start, end,
                            out.ctypes.data_as(ctypes.c_void_p))
    return out


def generate_graph_spec2012(scale, edgefactor=16, s1=2, s2=3, a_num=5500, bc_num=100):
    """generate_graph_spec2012 (mpi/benchmark_helper.hpp:316-322) without
    weights: 2^scale - 1 binary-tree tuples, then edgefactor * 2^scale R-MAT
    tuples at indices 2^scale - 1 onward."""
    n_init = (1 << scale) - 1
    init = np.zeros(n_init, EDGE_DTYPE)
    lib().or_initial_tree_edges(scale, s1, s2, 0, n_init, init.ctypes.data_as(ctypes.c_void_p))
    rmat = generate_edges(scale, edgefactor, s1, s2, a_num, bc_num, n_init, n_init + (edgefactor << scale))
    return np.concatenate([init, rmat])


def edge_weights(start, count, max_weight=0, s1=2, s2=3):
    """Weights of tuples [start, start+count) (generateWeight,
    generator/graph_generator.hpp:480-506): 0xBEEF when max_weight <= 0, else
    the per-edge draw % max_weight + 1."""
    out = np.zeros(count, np.int32)
    lib().or_edge_weights(s1, s2, start, count, max_weight, out.ctypes.data_as(ctypes.c_void_p))
    return out


def edge_endpoints(edges):
    v0 = edges["v0_low"].astype(np.int64) | ((edges["high"].astype(np.int64) & 0xFFFF) << 32)
    v1 = edges["v1_low"].astype(np.int64) | ((edges["high"].astype(np.int64) >> 16) << 32)
    return v0, v1


def pack_edges(v0, v1):
    v0 = np.asarray(v0, np.int64)
    v1 = np.asarray(v1, np.int64)
    e = np.zeros(len(v0), EDGE_DTYPE)
    e["v0_low"] = (v0 & 0xFFFFFFFF).astype(np.uint32)
    e["v1_low"] = (v1 & 0xFFFFFFFF).astype(np.uint32)
    e["high"] = (((v0 >> 32) & 0xFFFF) | ((v1 >> 16) & 0xFFFF0000)).astype(np.uint32)
    return e


class Csr:
    """Owning wrapper of or_csr (restates GraphConstructor2DCSR semantics)."""

    def __init__(self, edges):
        self.edges = np.ascontiguousarray(edges, EDGE_DTYPE)
        self._p = lib().or_csr_build(self.edges.ctypes.data_as(ctypes.c_void_p), len(self.edges))
        c = self._p.contents
        self.scale_estimate = c.scale_estimate
        self.num_verts = c.num_verts
        self.num_active = c.num_active
        self.num_self_loops = c.num_self_loops
        self.directed_before = c.directed_before
        self.nnz = c.nnz

    def __del__(self):
        if getattr(self, "_p", None):
            lib().or_csr_free(self._p)
            self._p = None

    def arrays(self):
        c = self._p.contents
        na, n = c.num_active, c.num_verts
        return {
            "offsets": np.ctypeslib.as_array(c.offsets, (na + 1,)).copy(),
            "cols": np.ctypeslib.as_array(c.cols, (max(c.nnz, 1),))[: c.nnz].copy(),
            "new2old": np.ctypeslib.as_array(c.new2old, (n,)).copy(),
            "old2new": np.ctypeslib.as_array(c.old2new, (n,)).copy(),
            "degree_raw": np.ctypeslib.as_array(c.degree_raw, (n,)).copy(),
            "self_loops": np.ctypeslib.as_array(c.self_loops, (n,)).copy(),
        }

    def has_edge(self, v):
        return bool(lib().or_csr_has_edge(self._p, v))

    def max_used_vertex(self):
        return lib().or_find_max_used_vertex(self._p)

    def find_roots(self, n, s1=2, s2=3):
        out = np.zeros(n, np.int64)
        k = lib().or_find_roots(self._p, s1, s2, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out[:k]

    def bfs(self, root, want_pred=True):
        """Returns (levels_by_orig uint8, pred int64 or None, trace dict)."""
        lv = np.zeros(self.num_verts, np.uint8)
        pred = np.zeros(self.num_verts, np.int64) if want_pred else None
        tr = OrBfsTrace()
        rc = lib().or_bfs(self._p, root, lv.ctypes.data_as(ctypes.c_void_p),
                          pred.ctypes.data_as(ctypes.c_void_p) if want_pred else None,
                          ctypes.byref(tr))
        if rc != 0:
            raise RuntimeError("oracle bfs failed rc=%d" % rc)
        L = tr.num_levels
        trace = {
            "num_levels": L,
            "direction": [tr.direction[i] for i in range(1, L)],
            "nq": [tr.nq[i] for i in range(1, L)],
            "bytes": [tr.bytes[i] for i in range(1, L)],
            "edges_examined": [tr.edges_examined[i] for i in range(1, L)],
            "total_bytes": tr.total_bytes,
            "reached": tr.reached,
            "teps_edges": tr.teps_edges,
        }
        return lv, pred, trace

    def msbfs_bytes(self, levels_new, pull_lanes=None, push_lanes=None):
        """pull_lanes / push_lanes: per-level lane masks (index = level) of a
        run with per-root directions; None = every lane pulls."""
        levels_new = np.ascontiguousarray(levels_new, np.uint8)
        lanes = levels_new.shape[0]
        nl = ctypes.c_int(0)
        per = np.zeros(OR_MAX_LEVELS + 2, np.float64)
        masks = []
        for m in (pull_lanes, push_lanes):
            if m is None:
                masks.append(None)
            else:
                a = np.zeros(OR_MAX_LEVELS + 2, np.uint64)
                m = np.asarray(m, np.uint64)
                a[: min(len(m), len(a))] = m[: len(a)]
                masks.append(a)
        ptrs = [None if m is None else m.ctypes.data_as(ctypes.c_void_p) for m in masks]
        tot = lib().or_msbfs_bytes(self._p, lanes, levels_new.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.byref(nl), per.ctypes.data_as(ctypes.c_void_p), ptrs[0], ptrs[1])
        return tot, per[: nl.value]


def validate(edges, nglobalverts, root, pred):
    """BfsValidation::validate restated. Returns (ok, edge_visit_count, message)."""
    edges = np.ascontiguousarray(edges, EDGE_DTYPE)
    pred = np.ascontiguousarray(pred, np.int64)
    cnt = ctypes.c_int64(0)
    msg = ctypes.create_string_buffer(512)
    ok = lib().or_validate(edges.ctypes.data_as(ctypes.c_void_p), len(edges), nglobalverts,
                           len(pred), root, pred.ctypes.data_as(ctypes.c_void_p),
                           ctypes.byref(cnt), msg, 512)
    return bool(ok), cnt.value, msg.value.decode()


def statistics(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(7, np.float64)
    lib().or_statistics(x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(x),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return dict(zip(["min", "q1", "median", "q3", "max", "mean", "std"], out))
