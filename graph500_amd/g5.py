"""Python host side over the C ABI (include/g5.h), mirroring the reference's
BfsBase operator interface (mpi/bfs.hpp:29-130):

    ctx = G5Context()                         # one GPU
    ctx.generate_kronecker(scale, 16)         # generate_graph_spec2010
    ctx.construct()                           # BfsBase::construct (VR=2)
    roots = ctx.find_roots(64)                # find_roots
    ctx.prepare_bfs()
    ctx.run_bfs(root)                         # timed region of mpi/main.cc:113-117
    pred = ctx.get_pred()                     # parent | level << 48, -1 unreached
    ctx.end_bfs()

The HIP library is required: importing works everywhere, but creating a
context without graph500_amd/lib/libg5.so or without a GPU raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libg5.so")

EDGE_DTYPE = np.dtype([("v0_low", "<u4"), ("v1_low", "<u4"), ("high", "<u4")])
WEIGHTED_EDGE_DTYPE = np.dtype([("v0_low", "<u4"), ("v1_low", "<u4"), ("high", "<u4"), ("weight", "<i4")])
MAX_LEVELS = 256

G5_OK = 0
ERRORS = {-1: "G5_ERR_ARG", -2: "G5_ERR_STATE", -3: "G5_ERR_HIP", -4: "G5_ERR_MEM",
          -5: "G5_ERR_LIMIT"}

# Every entry point of include/g5.h (the export test checks this list against
# the header and the library); include/g5_measure.h adds MEASURE_EXPORTS.
EXPORTS = [
    "g5_version", "g5_device_count", "g5_create", "g5_destroy", "g5_last_error",
    "g5_synchronize", "g5_edges_reserve", "g5_edges_append", "g5_edges_count",
    "g5_edges_download", "g5_edges_download_weighted", "g5_edges_release", "g5_generate_kronecker",
    "g5_generate_graph", "g5_construct",
    "g5_graph_info_get", "g5_has_edge", "g5_pred_size", "g5_find_roots", "g5_graph_download",
    "g5_prepare_bfs", "g5_run_bfs", "g5_get_pred", "g5_get_levels", "g5_teps_edges",
    "g5_bfs_stats_get", "g5_bfs_model_bytes", "g5_bfs_set_tuning", "g5_bfs_get_tuning", "g5_validate", "g5_validate_pred", "g5_validate_pred_noalloc", "g5_end_bfs",
    "g5_batch_prepare", "g5_batch_set_tuning", "g5_batch_get_tuning", "g5_run_bfs_batch",
    "g5_batch_get_levels", "g5_batch_teps_edges", "g5_batch_stats_get", "g5_batch_model_bytes",
    "g5_batch_level_lanes",
    "g5_batch_end", "g5_server_create", "g5_server_destroy", "g5_server_last_error", "g5_server_run",
    "g5_server_get_levels", "g5_server_release", "g5_server_graph_stats_get",
]
MEASURE_EXPORTS = ["g5_measure_gather"]


class G5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "G5_ERR"), code, msg))
        self.code = code


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("scale_estimate", ctypes.c_int32),
        ("reorder_mode", ctypes.c_int32),
        ("num_verts", ctypes.c_int64),
        ("num_active", ctypes.c_int64),
        ("num_tuples", ctypes.c_int64),
        ("num_self_loops", ctypes.c_int64),
        ("directed_before", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("max_degree", ctypes.c_int64),
        ("max_used_vertex", ctypes.c_int64),
        ("construct_ms", ctypes.c_double),
        ("generate_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BfsStats(ctypes.Structure):
    _fields_ = [
        ("num_levels", ctypes.c_int32),
        ("direction", ctypes.c_int32 * MAX_LEVELS),
        ("nq", ctypes.c_int64 * MAX_LEVELS),
        ("level_ms", ctypes.c_double * MAX_LEVELS),
        ("device_ms", ctypes.c_double),
        ("reached", ctypes.c_int64),
        ("bu_kernel", ctypes.c_int32 * MAX_LEVELS),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("num_roots", ctypes.c_int32),
        ("num_levels", ctypes.c_int32),
        ("direction", ctypes.c_int32 * MAX_LEVELS),
        ("pairs", ctypes.c_int64 * MAX_LEVELS),
        ("level_ms", ctypes.c_double * MAX_LEVELS),
        ("device_ms", ctypes.c_double),
        ("persistent_grid", ctypes.c_int32),
        ("persistent_levels", ctypes.c_int32),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [("answer_buffers", ctypes.c_int32), ("pull_filter", ctypes.c_int32), ("alpha", ctypes.c_double),
                ("min_admit", ctypes.c_int32), ("lane_rule", ctypes.c_int32), ("max_wait_s", ctypes.c_double),
                ("solo", ctypes.c_int32), ("solo_slots", ctypes.c_int32)]


class Query(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_int32), ("keep", ctypes.c_int32), ("root", ctypes.c_int64),
                ("arrival_s", ctypes.c_double)]


class QueryResult(ctypes.Structure):
    _fields_ = [("request", ctypes.c_int64), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("edges", ctypes.c_int64), ("levels", ctypes.c_int32), ("status", ctypes.c_int32)]


class ServerGraphStats(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int64), ("push_levels", ctypes.c_int64), ("pull_levels", ctypes.c_int64),
                ("mixed_levels", ctypes.c_int64), ("lane_levels", ctypes.c_int64), ("busy_ms", ctypes.c_double),
                ("completed", ctypes.c_int64), ("admitted_busy", ctypes.c_int64), ("admitted_idle", ctypes.c_int64),
                ("solo", ctypes.c_int64), ("solo_overlap", ctypes.c_int64)]


class ServerReport(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("completed", ctypes.c_int64), ("edges", ctypes.c_int64),
                ("levels_launched", ctypes.c_int64), ("max_active_lanes", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("admitted_busy", ctypes.c_int64), ("admit_events", ctypes.c_int64)]


QUERY_DTYPE = np.dtype([("graph", np.int32), ("keep", np.int32), ("root", np.int64), ("arrival_s", np.float64)])
RESULT_DTYPE = np.dtype([("request", np.int64), ("admit_s", np.float64), ("done_s", np.float64),
                         ("edges", np.int64), ("levels", np.int32), ("status", np.int32)])


class BatchTuning(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_double), ("pull_grid", ctypes.c_int64), ("sweep_below", ctypes.c_double),
                ("sweep_grid", ctypes.c_int64), ("persistent", ctypes.c_int32), ("persist_levels", ctypes.c_int32),
                ("persist_grid", ctypes.c_int64), ("persist_waves", ctypes.c_int32), ("pull_filter", ctypes.c_int32),
                ("beta", ctypes.c_double), ("filter_from", ctypes.c_double), ("pull_unroll", ctypes.c_int32),
                ("lane_rule", ctypes.c_int32), ("compact_below", ctypes.c_double), ("mixed_push", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


_lib = None


def load_library(path=LIB_PATH):
    """Loads libg5.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libg5.so not built: run `python -m graph500_amd.build` (%s)" % path)
    L = ctypes.CDLL(path)
    vp, i64, u64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
    sig = {
        "g5_version": ([], ctypes.c_char_p),
        "g5_device_count": ([ctypes.POINTER(i32)], i32),
        "g5_create": ([i32, ctypes.POINTER(vp)], i32),
        "g5_destroy": ([vp], i32),
        "g5_last_error": ([vp], ctypes.c_char_p),
        "g5_synchronize": ([vp], i32),
        "g5_edges_reserve": ([vp, i64], i32),
        "g5_edges_append": ([vp, vp, i64], i32),
        "g5_edges_count": ([vp, ctypes.POINTER(i64)], i32),
        "g5_edges_download": ([vp, i64, i64, vp], i32),
        "g5_edges_download_weighted": ([vp, i64, i64, i32, u64, u64, vp], i32),
        "g5_edges_release": ([vp], i32),
        "g5_generate_kronecker": ([vp, i32, i32, u64, u64, i32, i32], i32),
        "g5_generate_graph": ([vp, i32, i32, u64, u64, i32, i32, i32], i32),
        "g5_construct": ([vp, i32, i32], i32),
        "g5_graph_info_get": ([vp, ctypes.POINTER(GraphInfo)], i32),
        "g5_has_edge": ([vp, i64], i32),
        "g5_pred_size": ([vp, ctypes.POINTER(i64)], i32),
        "g5_find_roots": ([vp, u64, u64, i32, vp, ctypes.POINTER(i32)], i32),
        "g5_graph_download": ([vp, vp, vp, vp, vp], i32),
        "g5_prepare_bfs": ([vp], i32),
        "g5_run_bfs": ([vp, i64], i32),
        "g5_get_pred": ([vp, vp], i32),
        "g5_get_levels": ([vp, vp], i32),
        "g5_teps_edges": ([vp, ctypes.POINTER(i64)], i32),
        "g5_bfs_stats_get": ([vp, ctypes.POINTER(BfsStats)], i32),
        "g5_bfs_model_bytes": ([vp, vp, ctypes.POINTER(ctypes.c_double)], i32),
        "g5_validate": ([vp, vp, ctypes.POINTER(i32)], i32),
        "g5_bfs_set_tuning": ([vp, vp], i32),
        "g5_bfs_get_tuning": ([vp, vp], i32),
        "g5_validate_pred": ([vp, i64, vp, vp, ctypes.POINTER(i32), ctypes.POINTER(i64)], i32),
        "g5_validate_pred_noalloc": ([vp, i64, vp, vp, ctypes.POINTER(i32), ctypes.POINTER(i64)], i32),
        "g5_end_bfs": ([vp], i32),
        "g5_batch_prepare": ([vp], i32),
        "g5_batch_set_tuning": ([vp, ctypes.POINTER(BatchTuning)], i32),
        "g5_batch_get_tuning": ([vp, ctypes.POINTER(BatchTuning)], i32),
        "g5_run_bfs_batch": ([vp, vp, i32], i32),
        "g5_batch_get_levels": ([vp, i32, vp], i32),
        "g5_batch_teps_edges": ([vp, vp], i32),
        "g5_batch_stats_get": ([vp, ctypes.POINTER(BatchStats)], i32),
        "g5_batch_model_bytes": ([vp, vp, ctypes.POINTER(ctypes.c_double)], i32),
        "g5_batch_level_lanes": ([vp, vp, vp], i32),
        "g5_batch_end": ([vp], i32),
        "g5_server_create": ([vp, i32, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)], i32),
        "g5_server_destroy": ([vp], None),
        "g5_server_last_error": ([vp], ctypes.c_char_p),
        "g5_server_run": ([vp, vp, i64, ctypes.c_double, vp, ctypes.POINTER(ServerReport)], i32),
        "g5_server_get_levels": ([vp, i64, vp], i32),
        "g5_server_release": ([vp, i64], i32),
        "g5_server_graph_stats_get": ([vp, i32, ctypes.POINTER(ServerGraphStats)], i32),
        # include/g5_measure.h
        "g5_measure_gather": ([i32, i64, i64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class G5Context:
    """One GPU context; methods mirror BfsBase / EdgeListStorage."""

    def __init__(self, device=0):
        self._lib = load_library()
        n = ctypes.c_int(0)
        self._lib.g5_device_count(ctypes.byref(n))
        if n.value <= device:
            raise RuntimeError("no HIP device %d (found %d)" % (device, n.value))
        h = ctypes.c_void_p()
        rc = self._lib.g5_create(device, ctypes.byref(h))
        if rc != G5_OK:
            raise G5Error(rc, "g5_create failed")
        self._h = h

    def close(self):
        """Frees the context; raises G5Error (and keeps it) while a G5Server
        over it is still open."""
        if getattr(self, "_h", None):
            self._check(self._lib.g5_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except G5Error:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise G5Error(rc, self._lib.g5_last_error(self._h).decode())
        return rc

    # -- edge list ---------------------------------------------------------
    def generate_kronecker(self, scale, edgefactor=16, seed1=2, seed2=3, a_num=5700, bc_num=1900):
        self._check(self._lib.g5_generate_kronecker(self._h, scale, edgefactor, seed1, seed2,
                                                    a_num, bc_num))

    def generate_graph_spec2012(self, scale, edgefactor=16, seed1=2, seed2=3, a_num=5500, bc_num=100):
        """generate_graph_spec2012 without weights: binary-tree tuples first,
        then R-MAT tuples (initiator 5500/100) at their global indices."""
        self._check(self._lib.g5_generate_graph(self._h, scale, edgefactor, seed1, seed2, a_num, bc_num, 1))

    def edges_append(self, tuples):
        t = np.ascontiguousarray(tuples, EDGE_DTYPE)
        self._check(self._lib.g5_edges_append(self._h, _ptr(t), len(t)))

    def edges_count(self):
        n = ctypes.c_int64(0)
        self._check(self._lib.g5_edges_count(self._h, ctypes.byref(n)))
        return n.value

    def edges_download(self, start=0, count=None):
        if count is None:
            count = self.edges_count() - start
        out = np.zeros(count, EDGE_DTYPE)
        self._check(self._lib.g5_edges_download(self._h, start, count, _ptr(out)))
        return out

    def edges_download_weighted(self, start=0, count=None, max_weight=0, seed1=2, seed2=3):
        """Tuples [start, start+count) as WeightedEdge records (weights made
        on the GPU): 0xBEEF when max_weight <= 0, else the per-edge draw."""
        if count is None:
            count = self.edges_count() - start
        out = np.zeros(count, WEIGHTED_EDGE_DTYPE)
        self._check(self._lib.g5_edges_download_weighted(self._h, start, count, max_weight, seed1, seed2,
                                                         _ptr(out)))
        return out

    def edges_release(self):
        self._check(self._lib.g5_edges_release(self._h))

    # -- graph -------------------------------------------------------------
    def construct(self, reorder_mode=2, release_edges=False):
        self._check(self._lib.g5_construct(self._h, reorder_mode, 1 if release_edges else 0))

    def graph_info(self):
        gi = GraphInfo()
        self._check(self._lib.g5_graph_info_get(self._h, ctypes.byref(gi)))
        return gi.as_dict()

    def has_edge(self, v):
        return bool(self._check(self._lib.g5_has_edge(self._h, v)))

    def pred_size(self):
        n = ctypes.c_int64(0)
        self._check(self._lib.g5_pred_size(self._h, ctypes.byref(n)))
        return n.value

    def find_roots(self, num_roots, seed1=2, seed2=3):
        out = np.zeros(max(num_roots, 1), np.int64)
        k = ctypes.c_int(0)
        self._check(self._lib.g5_find_roots(self._h, seed1, seed2, num_roots, _ptr(out),
                                            ctypes.byref(k)))
        return out[: k.value]

    def graph_download(self):
        gi = self.graph_info()
        rows = gi["num_verts"] if gi["reorder_mode"] == 0 else gi["num_active"]
        offsets = np.zeros(rows + 1, np.uint64)
        cols = np.zeros(max(gi["nnz"], 1), np.uint32)
        new2old = np.zeros(gi["num_verts"], np.uint32)
        weight = np.zeros(max(rows, 1), np.uint32)
        self._check(self._lib.g5_graph_download(self._h, _ptr(offsets), _ptr(cols), _ptr(new2old),
                                                _ptr(weight)))
        return {"offsets": offsets, "cols": cols[: gi["nnz"]], "new2old": new2old,
                "weight": weight[:rows]}

    # -- BFS ---------------------------------------------------------------
    def prepare_bfs(self):
        self._check(self._lib.g5_prepare_bfs(self._h))

    def run_bfs(self, root):
        self._check(self._lib.g5_run_bfs(self._h, int(root)))

    def get_pred(self, out=None):
        if out is None:
            out = np.zeros(self.pred_size(), np.int64)
        self._check(self._lib.g5_get_pred(self._h, _ptr(out)))
        return out

    def get_levels(self, out=None):
        if out is None:
            out = np.zeros(self.pred_size(), np.uint8)
        self._check(self._lib.g5_get_levels(self._h, _ptr(out)))
        return out

    def teps_edges(self):
        n = ctypes.c_int64(0)
        self._check(self._lib.g5_teps_edges(self._h, ctypes.byref(n)))
        return n.value

    def bfs_stats(self):
        st = BfsStats()
        self._check(self._lib.g5_bfs_stats_get(self._h, ctypes.byref(st)))
        L = st.num_levels
        return {
            "num_levels": L,
            "direction": [st.direction[i] for i in range(1, L)],
            "nq": [st.nq[i] for i in range(1, L)],
            "level_ms": [st.level_ms[i] for i in range(1, L)],
            "device_ms": st.device_ms,
            "reached": st.reached,
            "bu_kernel": [st.bu_kernel[i] for i in range(1, L)],
        }

    def model_bytes(self):
        """(per-level algorithmic bytes [level 1..], total) of the last run."""
        per = np.zeros(MAX_LEVELS, np.float64)
        tot = ctypes.c_double(0)
        self._check(self._lib.g5_bfs_model_bytes(self._h, _ptr(per), ctypes.byref(tot)))
        L = self.bfs_stats()["num_levels"]
        return per[1:L].tolist(), tot.value

    BFS_TUNING = ("bin_edges", "td_pieces", "block_shift", "td_head", "seg_max", "small_edges", "tail_pull")

    def bfs_set_tuning(self, bin_edges=None, td_pieces=None, block_shift=None, td_head=None, seg_max=None,
                       small_edges=None, tail_pull=None):
        """Single-root tuning; a knob left None keeps its value.  bin_edges:
        top-down levels with at least this many frontier edges are binned by
        target range (0 = never).  td_pieces: a growing bulk top-down level
        takes its rows in this many pieces, piece index first (0 = row order).
        block_shift: a growing bulk top-down level takes its edges by target
        block of 2^block_shift ids (0 = off; it then falls back to td_pieces).
        td_head: a growing bulk top-down level with E frontier edges settles
        its targets of degree >= td_head / 16 * (nnz / E)^2 (a prefix of the
        degree-ordered rows) with a bottom-up probe and skips their edges
        (0 = off; degree-ordered graphs only).  seg_max: most segments of a
        segment-list level (0 = 2^24 - 1); a level beyond it is taken in row
        order (tests lower it to reach that fallback on small graphs).
        small_edges: consecutive top-down levels with at most 4096 frontier
        vertices and this many frontier edges run in one launch that applies
        the direction rule on the device (0 = off).  tail_pull: a top-down
        level after the bottom-up phase whose frontier holds more than 4096
        and 1/tail_pull as many vertices as remain unvisited is found by the
        bottom-up kernel, keeping its top-down place in the schedule
        (default 4; 0 = off; at most 64)."""
        t = np.zeros(len(self.BFS_TUNING), np.int64)
        self._check(self._lib.g5_bfs_get_tuning(self._h, _ptr(t)))
        for i, v in enumerate((bin_edges, td_pieces, block_shift, td_head, seg_max, small_edges, tail_pull)):
            if v is not None:
                t[i] = v
        self._check(self._lib.g5_bfs_set_tuning(self._h, _ptr(t)))

    def bfs_tuning(self):
        t = np.zeros(len(self.BFS_TUNING), np.int64)
        self._check(self._lib.g5_bfs_get_tuning(self._h, _ptr(t)))
        return {k: int(v) for k, v in zip(self.BFS_TUNING, t)}

    VALIDATION_RULES = ["root", "self_parent", "depth", "edge", "parent_edge", "range"]

    def validate(self):
        """(valid, {rule: violations}) of the last run, checked on the GPU."""
        err = np.zeros(6, np.int64)
        ok = ctypes.c_int(0)
        self._check(self._lib.g5_validate(self._h, _ptr(err), ctypes.byref(ok)))
        return bool(ok.value), dict(zip(self.VALIDATION_RULES, err.tolist()))

    def validate_pred(self, root, pred):
        """(valid, {rule: violations}, traversed edges) of a pred array of
        pred_size() int64 entries (parent | level << 48, -1 unreached),
        checked on the GPU against the resident graph."""
        pred = np.ascontiguousarray(pred, dtype=np.int64)
        if pred.shape != (self.pred_size(),):
            raise ValueError("pred must hold pred_size() = %d entries" % self.pred_size())
        err = np.zeros(6, np.int64)
        ok = ctypes.c_int(0)
        visited = ctypes.c_int64(0)
        self._check(self._lib.g5_validate_pred(self._h, int(root), _ptr(pred), _ptr(err), ctypes.byref(ok),
                                               ctypes.byref(visited)))
        return bool(ok.value), dict(zip(self.VALIDATION_RULES, err.tolist())), int(visited.value)

    def validate_pred_noalloc(self, root, pred):
        """validate_pred through g5_validate_pred_noalloc (no device allocation;
        graphs of up to about 2^22 vertex ids)."""
        pred = np.ascontiguousarray(pred, dtype=np.int64)
        if pred.shape != (self.pred_size(),):
            raise ValueError("pred must hold pred_size() = %d entries" % self.pred_size())
        err = np.zeros(6, np.int64)
        ok = ctypes.c_int(0)
        visited = ctypes.c_int64(0)
        self._check(self._lib.g5_validate_pred_noalloc(self._h, int(root), _ptr(pred), _ptr(err), ctypes.byref(ok),
                                                       ctypes.byref(visited)))
        return bool(ok.value), dict(zip(self.VALIDATION_RULES, err.tolist())), int(visited.value)

    def end_bfs(self):
        self._check(self._lib.g5_end_bfs(self._h))

    # -- batched BFS (up to 64 roots per sweep) ---------------------------------
    def batch_prepare(self):
        self._check(self._lib.g5_batch_prepare(self._h))

    TUNING_FIELDS = ("alpha", "pull_grid", "sweep_below", "sweep_grid", "persistent", "persist_levels",
                     "persist_grid", "persist_waves", "pull_filter", "beta", "filter_from", "pull_unroll",
                     "lane_rule", "compact_below", "mixed_push")

    def batch_set_tuning(self, **kw):
        """Sets any of TUNING_FIELDS; the others keep their values."""
        bad = set(kw) - set(self.TUNING_FIELDS)
        if bad:
            raise TypeError("unknown tuning fields %s" % sorted(bad))
        t = BatchTuning()
        self._check(self._lib.g5_batch_get_tuning(self._h, ctypes.byref(t)))
        for name, val in kw.items():
            if val is not None:
                setattr(t, name, val)
        self._check(self._lib.g5_batch_set_tuning(self._h, ctypes.byref(t)))

    def batch_tuning(self):
        t = BatchTuning()
        self._check(self._lib.g5_batch_get_tuning(self._h, ctypes.byref(t)))
        return {name: getattr(t, name) for name in self.TUNING_FIELDS}

    def run_bfs_batch(self, roots):
        r = np.ascontiguousarray(roots, np.int64)
        self._check(self._lib.g5_run_bfs_batch(self._h, _ptr(r), len(r)))

    def batch_levels(self, lane, out=None):
        if out is None:
            out = np.zeros(self.pred_size(), np.uint8)
        self._check(self._lib.g5_batch_get_levels(self._h, lane, _ptr(out)))
        return out

    def batch_teps_edges(self):
        st = self.batch_stats()
        out = np.zeros(st["num_roots"], np.int64)
        self._check(self._lib.g5_batch_teps_edges(self._h, _ptr(out)))
        return out

    def batch_stats(self):
        st = BatchStats()
        self._check(self._lib.g5_batch_stats_get(self._h, ctypes.byref(st)))
        L = st.num_levels
        return {
            "num_roots": st.num_roots,
            "num_levels": L,
            "direction": [st.direction[i] for i in range(1, L)],
            "pairs": [st.pairs[i] for i in range(1, L)],
            "level_ms": [st.level_ms[i] for i in range(1, L)],
            "device_ms": st.device_ms,
            "persistent_grid": st.persistent_grid,
            "persistent_levels": st.persistent_levels,
        }

    def batch_model_bytes(self):
        per = np.zeros(MAX_LEVELS, np.float64)
        tot = ctypes.c_double(0)
        self._check(self._lib.g5_batch_model_bytes(self._h, _ptr(per), ctypes.byref(tot)))
        L = self.batch_stats()["num_levels"]
        return per[1:L].tolist(), tot.value

    def batch_level_lanes(self):
        """(push, pull): per-level lane masks of the last batch (index = level)."""
        push = np.zeros(MAX_LEVELS, np.uint64)
        pull = np.zeros(MAX_LEVELS, np.uint64)
        self._check(self._lib.g5_batch_level_lanes(self._h, _ptr(push), _ptr(pull)))
        return push, pull

    def batch_end(self):
        self._check(self._lib.g5_batch_end(self._h))

    def version(self):
        """g5_version(): library build and GPU architecture."""
        return self._lib.g5_version().decode()

    def synchronize(self):
        self._check(self._lib.g5_synchronize(self._h))


class G5Server:
    """Serving loop over graphs resident in HBM (g5_server_*): continuous
    batching of (graph, root) BFS queries, up to 64 in flight per graph."""

    def __init__(self, contexts, answer_buffers=128, alpha=64.0, pull_filter=0, min_admit=64, max_wait_s=0.3,
                 lane_rule=3, solo=1, solo_slots=1):
        """Defaults: a graph with searches in flight admits waiting queries
        into its free lanes at a level boundary while every search in flight
        still pushes (newcomers join a young cohort), once 64 of them can
        start together, or once the oldest has waited 0.3 s; each search
        pulls while its frontier's edges exceed 1/64 of the edges it has not
        reached, and pushes otherwise.  solo k (1..64): a graph with no search
        in flight and at most k waiting queries answers the oldest with the
        single-root kernels, then looks again (the contexts are prepared for
        g5_run_bfs if they are not); 0: every query takes a lane.
        solo_slots: single-root runs of one graph that may go on at once
        (each slot beyond the first holds a BFS state of its own)."""
        self._lib = load_library()
        self._ctxs = list(contexts)  # the server uses their graphs and streams
        arr = (ctypes.c_void_p * len(self._ctxs))(*[c._h for c in self._ctxs])
        cfg = ServerConfig(answer_buffers, pull_filter, alpha, min_admit, lane_rule, max_wait_s, solo, solo_slots)
        h = ctypes.c_void_p()
        rc = self._lib.g5_server_create(arr, len(self._ctxs), ctypes.byref(cfg), ctypes.byref(h))
        if rc != G5_OK:
            raise G5Error(rc, "g5_server_create failed (graphs built with reorder mode 1 or 2 on one device?)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.g5_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise G5Error(rc, self._lib.g5_server_last_error(self._h).decode())
        return rc

    def run(self, queries, timeout_s=600.0):
        """queries: array of QUERY_DTYPE (graph, keep, root, arrival_s),
        arrivals non-decreasing.  Returns (results array of RESULT_DTYPE,
        report dict)."""
        q = np.ascontiguousarray(queries, dtype=QUERY_DTYPE)
        res = np.zeros(len(q), RESULT_DTYPE)
        rep = ServerReport()
        self._check(self._lib.g5_server_run(self._h, _ptr(q), len(q), float(timeout_s), _ptr(res),
                                            ctypes.byref(rep)))
        return res, {"wall_s": rep.wall_s, "completed": rep.completed, "edges": rep.edges,
                     "levels_launched": rep.levels_launched, "max_active_lanes": rep.max_active_lanes,
                     "admitted_busy": rep.admitted_busy, "admit_events": rep.admit_events}

    def levels(self, request, n):
        out = np.zeros(n, np.uint8)
        self._check(self._lib.g5_server_get_levels(self._h, int(request), _ptr(out)))
        return out

    def release(self, request):
        self._check(self._lib.g5_server_release(self._h, int(request)))

    def graph_stats(self, graph):
        st = ServerGraphStats()
        self._check(self._lib.g5_server_graph_stats_get(self._h, int(graph), ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerGraphStats._fields_}
