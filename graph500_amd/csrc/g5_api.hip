// g5_api.hip -- the extern "C" boundary (include/g5.h).  Every entry point
// catches g5::Error and reports a G5_ERR_* code plus a message; nothing
// throws across the ABI.
#include "g5_ctx.hpp"

#include <string.h>

#include <new>

using namespace g5;

namespace {

template <class F>
int guarded(g5_ctx* ctx, F&& f) {
  if (!ctx) return G5_ERR_ARG;
  try {
    ctx->err.clear();
    f();
    return G5_OK;
  } catch (const Error& e) {
    ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return G5_ERR_MEM;
  } catch (...) {
    ctx->err = "unexpected exception";
    return G5_ERR_STATE;
  }
}

void set_device(g5_ctx* ctx) { G5_HIP(hipSetDevice(ctx->device)); }

}  // namespace

extern "C" {

const char* g5_version(void) { return "g5-mi355x 0.1 (gfx950)"; }

int g5_device_count(int* count) {
  if (!count) return G5_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return G5_OK;
}

int g5_create(int device, g5_ctx** out) {
  if (!out) return G5_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return G5_ERR_ARG;
  g5_ctx* ctx = new (std::nothrow) g5_ctx();
  if (!ctx) return G5_ERR_MEM;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return G5_ERR_HIP;
  }
  *out = ctx;
  return G5_OK;
}

int g5_destroy(g5_ctx* ctx) {
  if (!ctx) return G5_OK;
  if (ctx->servers) {
    ctx->err = "g5_destroy: a server still serves this context's graph (g5_server_destroy first)";
    return G5_ERR_STATE;
  }
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->bfs.release();
  ctx->ms.release();
  ctx->graph.clear();
  ctx->edges.release();
  ctx->wstage.release();
  ctx->acc.release();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return G5_OK;
}

const char* g5_last_error(const g5_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int g5_synchronize(g5_ctx* ctx) {
  return guarded(ctx, [&] {
    set_device(ctx);
    G5_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int g5_edges_reserve(g5_ctx* ctx, int64_t capacity) {
  return guarded(ctx, [&] {
    G5_CHECK(capacity >= 0, "g5_edges_reserve: negative capacity");
    set_device(ctx);
    if ((size_t)capacity <= ctx->edges.n) return;
    DevBuf<g5_packed_edge> nb;
    nb.alloc(capacity);
    if (ctx->num_edges)
      G5_HIP(hipMemcpyAsync(nb.p, ctx->edges.p, ctx->num_edges * sizeof(g5_packed_edge),
                            hipMemcpyDeviceToDevice, ctx->stream));
    G5_HIP(hipStreamSynchronize(ctx->stream));
    std::swap(nb.p, ctx->edges.p);
    std::swap(nb.n, ctx->edges.n);
  });
}

int g5_edges_append(g5_ctx* ctx, const g5_packed_edge* tuples, int64_t count) {
  return guarded(ctx, [&] {
    G5_CHECK(count >= 0 && (tuples || count == 0), "g5_edges_append: bad arguments");
    if (count == 0) return;
    set_device(ctx);
    const int64_t need = ctx->num_edges + count;
    if ((size_t)need > ctx->edges.n) {
      // grow geometrically like EdgeListStorage's 3% slack, at least 'need'
      int64_t cap = need + need / 2;
      if (g5_edges_reserve(ctx, cap) != G5_OK) throw Error{ctx->err, G5_ERR_MEM};
    }
    G5_HIP(hipMemcpyAsync(ctx->edges.p + ctx->num_edges, tuples, count * sizeof(g5_packed_edge),
                          hipMemcpyHostToDevice, ctx->stream));
    G5_HIP(hipStreamSynchronize(ctx->stream));
    ctx->num_edges = need;
  });
}

int g5_edges_count(g5_ctx* ctx, int64_t* count) {
  return guarded(ctx, [&] {
    G5_CHECK(count, "g5_edges_count: null output");
    *count = ctx->num_edges;
  });
}

int g5_edges_download(g5_ctx* ctx, int64_t start, int64_t count, g5_packed_edge* out) {
  return guarded(ctx, [&] {
    G5_CHECK(out || count == 0, "g5_edges_download: null output");
    G5_CHECK(start >= 0 && count >= 0 && start + count <= ctx->num_edges,
             "g5_edges_download: range outside the edge list");
    if (!count) return;
    set_device(ctx);
    G5_HIP(hipMemcpyAsync(out, ctx->edges.p + start, count * sizeof(g5_packed_edge),
                          hipMemcpyDeviceToHost, ctx->stream));
    G5_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int g5_edges_download_weighted(g5_ctx* ctx, int64_t start, int64_t count, int max_weight, uint64_t seed1,
                               uint64_t seed2, g5_weighted_edge* out) {
  return guarded(ctx, [&] {
    G5_CHECK(out || count == 0, "g5_edges_download_weighted: null output");
    G5_CHECK(start >= 0 && count >= 0 && start + count <= ctx->num_edges,
             "g5_edges_download_weighted: range outside the edge list");
    if (!count) return;
    set_device(ctx);
    // through a device staging buffer, 8 Mi tuples at a time (the chunk of
    // EdgeListStorage, generator/graph_generator.hpp:35)
    const int64_t kChunk = INT64_C(1) << 23;
    ctx->wstage.alloc((size_t)std::min(count, kChunk));
    for (int64_t off = 0; off < count; off += kChunk) {
      const int64_t n = std::min(kChunk, count - off);
      launch_weighted(ctx->device, ctx->edges.p + start + off, start + off, n, max_weight, seed1, seed2,
                      ctx->wstage.p, ctx->stream);
      G5_HIP(hipMemcpyAsync(out + off, ctx->wstage.p, n * sizeof(g5_weighted_edge), hipMemcpyDeviceToHost,
                            ctx->stream));
      G5_HIP(hipStreamSynchronize(ctx->stream));
    }
  });
}

int g5_edges_release(g5_ctx* ctx) {
  return guarded(ctx, [&] {
    set_device(ctx);
    G5_HIP(hipStreamSynchronize(ctx->stream));
    ctx->edges.release();
    ctx->num_edges = 0;
  });
}

int g5_generate_graph(g5_ctx* ctx, int scale, int edgefactor, uint64_t seed1, uint64_t seed2, int a_num,
                      int bc_num, int initial_edges) {
  return guarded(ctx, [&] {
    G5_CHECK(edgefactor >= 1, "g5_generate_kronecker: edgefactor must be >= 1");
    G5_CHECK(scale >= 1 && scale <= 40, "g5_generate_kronecker: scale must be in [1, 40]");
    G5_CHECK(initial_edges == 0 || initial_edges == 1, "g5_generate_graph: initial_edges must be 0 or 1");
    set_device(ctx);
    const int64_t n_init = initial_edges == 1 ? ((int64_t)1 << scale) - 1 : 0;
    const int64_t m = ((int64_t)1 << scale) * edgefactor + n_init;
    ctx->num_edges = 0;
    if ((size_t)m > ctx->edges.n) {
      ctx->edges.release();
      ctx->edges.alloc(m);
    }
    hipEvent_t t0, t1;
    G5_HIP(hipEventCreate(&t0));
    G5_HIP(hipEventCreate(&t1));
    G5_HIP(hipEventRecord(t0, ctx->stream));
    // the initial tuples come first; the R-MAT tuples keep their global indices
    launch_initial_tree(ctx->edges.p, n_init, scale, seed1, seed2, ctx->stream);
    launch_kronecker(ctx->device, ctx->edges.p + n_init, n_init, m - n_init, scale, seed1, seed2, a_num, bc_num,
                     ctx->stream);
    G5_HIP(hipEventRecord(t1, ctx->stream));
    G5_HIP(hipEventSynchronize(t1));
    float ms = 0;
    G5_HIP(hipEventElapsedTime(&ms, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    ctx->generate_ms = ms;
    ctx->num_edges = m;
  });
}

int g5_generate_kronecker(g5_ctx* ctx, int scale, int edgefactor, uint64_t seed1, uint64_t seed2,
                          int a_num, int bc_num) {
  return g5_generate_graph(ctx, scale, edgefactor, seed1, seed2, a_num, bc_num, 0);
}

int g5_construct(g5_ctx* ctx, int reorder_mode, int release_edges) {
  return guarded(ctx, [&] {
    G5_CHECK_CODE(!ctx->servers, G5_ERR_STATE,
                  "g5_construct: a server serves this context's graph (g5_server_destroy first)");
    set_device(ctx);
    ctx->bfs.release();
    ctx->ms.release();
    build_graph(ctx, reorder_mode, release_edges != 0);
  });
}

int g5_graph_info_get(g5_ctx* ctx, g5_graph_info* out) {
  return guarded(ctx, [&] {
    G5_CHECK(out, "g5_graph_info_get: null output");
    const Graph& g = ctx->graph;
    if (!g.built) throw Error{"g5_graph_info_get: no graph (call g5_construct first)", G5_ERR_STATE};
    memset(out, 0, sizeof(*out));
    out->scale_estimate = g.scale;
    out->reorder_mode = g.reorder_mode;
    out->num_verts = g.n;
    out->num_active = g.n_active;
    out->num_tuples = g.num_tuples;
    out->num_self_loops = g.self_loops;
    out->directed_before = g.directed_before;
    out->nnz = g.nnz;
    out->max_degree = g.max_degree;
    out->max_used_vertex = g.max_used_vertex;
    out->construct_ms = g.construct_ms;
    out->generate_ms = ctx->generate_ms;
  });
}

int g5_has_edge(g5_ctx* ctx, int64_t v) {
  int r = 0;
  int rc = guarded(ctx, [&] {
    const Graph& g = ctx->graph;
    if (!g.built) throw Error{"g5_has_edge: no graph", G5_ERR_STATE};
    if (v < 0 || v >= g.n) {
      r = 0;
      return;
    }
    r = (int)((g.has_edge_words[v >> 6] >> (v & 63)) & 1);
  });
  return rc == G5_OK ? r : rc;
}

int g5_pred_size(g5_ctx* ctx, int64_t* size) {
  return guarded(ctx, [&] {
    G5_CHECK(size, "g5_pred_size: null output");
    if (!ctx->graph.built) throw Error{"g5_pred_size: no graph", G5_ERR_STATE};
    *size = ctx->graph.n;
  });
}

// find_roots (mpi/benchmark_helper.hpp:475-508): candidate i is
// floor((d0 + d1) * N) mod N from the two doubles at stream position
// 2*counter of the (seed1, seed2) MRG skipped by 2^129 (make_random_numbers,
// generator/graph_generator.hpp:338-354); duplicates and vertices without
// edges are skipped; the search gives up after 2N draws.
int g5_find_roots(g5_ctx* ctx, uint64_t seed1, uint64_t seed2, int num_roots, int64_t* roots,
                  int* num_found) {
  return guarded(ctx, [&] {
    G5_CHECK(roots && num_found && num_roots >= 0, "g5_find_roots: bad arguments");
    const Graph& g = ctx->graph;
    if (!g.built) throw Error{"g5_find_roots: no graph", G5_ERR_STATE};
    const int64_t N = g.n;
    int64_t counter = 0;
    int idx;
    for (idx = 0; idx < num_roots; ++idx) {
      int64_t root;
      for (;;) {
        MrgState st;
        host_make_mrg_seed(seed1, seed2, &st);
        host_mrg_skip(&st, 2, 0, 2 * (uint64_t)counter);
        const double d0 = host_mrg_get_double(&st);
        const double d1 = host_mrg_get_double(&st);
        volatile double sum = d0 + d1;
        volatile double scaled = sum * (double)N;
        root = (int64_t)scaled % N;
        counter += 2;
        if (counter > 2 * N) break;
        bool dup = false;
        for (int i = 0; i < idx; ++i)
          if (roots[i] == root) {
            dup = true;
            break;
          }
        if (dup) continue;
        if ((g.has_edge_words[root >> 6] >> (root & 63)) & 1) break;
      }
      roots[idx] = root;
    }
    *num_found = idx;
  });
}

int g5_graph_download(g5_ctx* ctx, uint64_t* offsets, uint32_t* cols, uint32_t* new2old,
                      uint32_t* weight) {
  return guarded(ctx, [&] {
    const Graph& g = ctx->graph;
    if (!g.built) throw Error{"g5_graph_download: no graph", G5_ERR_STATE};
    set_device(ctx);
    hipStream_t s = ctx->stream;
    if (offsets)
      G5_HIP(hipMemcpyAsync(offsets, g.offsets.p, (g.n_rows + 1) * 8, hipMemcpyDeviceToHost, s));
    if (cols && g.nnz) G5_HIP(hipMemcpyAsync(cols, g.cols.p, g.nnz * 4, hipMemcpyDeviceToHost, s));
    if (new2old) G5_HIP(hipMemcpyAsync(new2old, g.new2old.p, g.n * 4, hipMemcpyDeviceToHost, s));
    if (weight) G5_HIP(hipMemcpyAsync(weight, g.weight.p, g.n_rows * 4, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
  });
}

int g5_prepare_bfs(g5_ctx* ctx) {
  return guarded(ctx, [&] {
    set_device(ctx);
    bfs_prepare(ctx);
  });
}

int g5_run_bfs(g5_ctx* ctx, int64_t root) {
  return guarded(ctx, [&] {
    set_device(ctx);
    bfs_run(ctx, root);
  });
}

int g5_get_pred(g5_ctx* ctx, int64_t* pred) {
  return guarded(ctx, [&] {
    G5_CHECK(pred, "g5_get_pred: null output");
    set_device(ctx);
    bfs_get_pred(ctx, pred);
  });
}

int g5_get_levels(g5_ctx* ctx, uint8_t* levels) {
  return guarded(ctx, [&] {
    G5_CHECK(levels, "g5_get_levels: null output");
    set_device(ctx);
    bfs_get_levels(ctx, levels);
  });
}

int g5_teps_edges(g5_ctx* ctx, int64_t* count) {
  return guarded(ctx, [&] {
    G5_CHECK(count, "g5_teps_edges: null output");
    set_device(ctx);
    *count = bfs_teps_edges(ctx);
  });
}

int g5_bfs_stats_get(g5_ctx* ctx, g5_bfs_stats* out) {
  return guarded(ctx, [&] {
    G5_CHECK(out, "g5_bfs_stats_get: null output");
    if (!ctx->bfs.have_result) throw Error{"g5_bfs_stats_get: no BFS result", G5_ERR_STATE};
    *out = ctx->bfs.stats;
  });
}

int g5_bfs_model_bytes(g5_ctx* ctx, double* per_level, double* total) {
  return guarded(ctx, [&] {
    set_device(ctx);
    bfs_model_bytes(ctx, per_level, total);
  });
}

int g5_validate(g5_ctx* ctx, int64_t* errors, int* valid) {
  return guarded(ctx, [&] {
    G5_CHECK(errors && valid, "g5_validate: null output");
    set_device(ctx);
    bfs_validate(ctx, errors);
    *valid = 1;
    for (int i = 0; i < 6; ++i)
      if (errors[i]) *valid = 0;
  });
}

int g5_validate_pred(g5_ctx* ctx, int64_t root, const int64_t* pred, int64_t* errors, int* valid,
                     int64_t* edge_visit_count) {
  return guarded(ctx, [&] {
    G5_CHECK(pred && errors && valid, "g5_validate_pred: null argument");
    set_device(ctx);
    const int64_t visited = pred_validate(ctx, root, pred, errors);
    *valid = 1;
    for (int i = 0; i < 6; ++i)
      if (errors[i]) *valid = 0;
    if (edge_visit_count) *edge_visit_count = visited;
  });
}

int g5_validate_pred_noalloc(g5_ctx* ctx, int64_t root, const int64_t* pred, int64_t* errors, int* valid,
                             int64_t* edge_visit_count) {
  return guarded(ctx, [&] {
    G5_CHECK(pred && errors && valid, "g5_validate_pred_noalloc: null argument");
    set_device(ctx);
    const int64_t visited = pred_validate_noalloc(ctx, root, pred, errors);
    *valid = 1;
    for (int i = 0; i < 6; ++i)
      if (errors[i]) *valid = 0;
    if (edge_visit_count) *edge_visit_count = visited;
  });
}

int g5_bfs_set_tuning(g5_ctx* ctx, const g5_bfs_tuning* t) {
  return guarded(ctx, [&] {
    G5_CHECK(t && t->bin_edges >= 0 && t->td_pieces >= 0 && t->td_pieces <= 4096 &&
             (t->block_shift == 0 || (t->block_shift >= 12 && t->block_shift <= 31)) && t->td_head >= 0 &&
             t->seg_max >= 0 && t->seg_max < (INT64_C(1) << 24) && t->small_edges >= 0 &&
             t->small_edges <= (INT64_C(1) << 20) && t->tail_pull >= 0 && t->tail_pull <= 64,
             "g5_bfs_set_tuning: bad tuning");
    ctx->bfs.bin_edges = t->bin_edges;
    ctx->bfs.td_pieces = t->td_pieces;
    ctx->bfs.hub_shift = t->block_shift;
    ctx->bfs.td_head = t->td_head;
    ctx->bfs.seg_max = t->seg_max;
    ctx->bfs.small_edges = t->small_edges;
    ctx->bfs.tail_pull = t->tail_pull;
    if (ctx->bfs.prepared) {
      // segment buffers and block index for the new tuning, outside any run
      set_device(ctx);
      bfs_tuning_buffers(ctx);
    }
  });
}

int g5_bfs_get_tuning(g5_ctx* ctx, g5_bfs_tuning* t) {
  return guarded(ctx, [&] {
    G5_CHECK(t, "g5_bfs_get_tuning: null output");
    memset(t, 0, sizeof(*t));
    t->bin_edges = ctx->bfs.bin_edges;
    t->td_pieces = ctx->bfs.td_pieces;
    t->block_shift = ctx->bfs.hub_shift;
    t->td_head = ctx->bfs.td_head;
    t->seg_max = ctx->bfs.seg_max;
    t->small_edges = ctx->bfs.small_edges;
    t->tail_pull = ctx->bfs.tail_pull;
  });
}

int g5_end_bfs(g5_ctx* ctx) {
  return guarded(ctx, [&] {
    set_device(ctx);
    G5_HIP(hipStreamSynchronize(ctx->stream));
    ctx->bfs.release();
  });
}

int g5_batch_prepare(g5_ctx* ctx) {
  return guarded(ctx, [&] {
    set_device(ctx);
    ms_prepare(ctx);
  });
}

int g5_batch_set_tuning(g5_ctx* ctx, const g5_batch_tuning* t) {
  return guarded(ctx, [&] {
    G5_CHECK(t && t->alpha > 0 && t->pull_grid >= 1 && t->pull_grid <= (1 << 20) && t->sweep_below >= 0 &&
                 t->sweep_grid >= 1 && t->sweep_grid <= (1 << 20) && (t->persistent == 0 || t->persistent == 1) &&
                 t->persist_levels >= 2 && t->persist_levels <= G5_MAX_LEVELS - 2 && t->persist_grid >= 0 &&
                 t->persist_grid <= (1 << 20) &&
                 (t->persist_waves == 4 || t->persist_waves == 6 || t->persist_waves == 8) &&
                 t->pull_filter >= 0 && t->pull_filter <= 2 && t->beta >= 0 && t->compact_below >= 0 &&
                 t->filter_from >= 0 && t->filter_from <= 1 && t->pull_unroll >= 0 && t->pull_unroll <= 7 &&
                 (t->lane_rule == 0 || t->lane_rule == 1) && (t->mixed_push == 0 || t->mixed_push == 1) &&
                 t->reserved == 0,
             "g5_batch_set_tuning: bad tuning");
    ctx->ms.persistent = t->persistent;
    ctx->ms.persist_levels = t->persist_levels;
    ctx->ms.persist_grid = t->persist_grid;
    ctx->ms.persist_waves = t->persist_waves;
    ctx->ms.pull_filter = t->pull_filter;
    ctx->ms.beta = t->beta;
    ctx->ms.compact_below = t->compact_below;
    ctx->ms.filter_from = t->filter_from;
    ctx->ms.pull_unroll = t->pull_unroll;
    ctx->ms.lane_rule = t->lane_rule;
    ctx->ms.mixed_push = t->mixed_push;
    ctx->ms.alpha = t->alpha;
    ctx->ms.pull_grid = t->pull_grid;
    ctx->ms.sweep_below = t->sweep_below;
    ctx->ms.sweep_grid = t->sweep_grid;
    if (ctx->ms.prepared && t->compact_below > 0) {
      // the compact frontier's buffers, now rather than inside a batch
      set_device(ctx);
      ms_alloc_compact(ctx);
    }
  });
}

int g5_batch_get_tuning(g5_ctx* ctx, g5_batch_tuning* t) {
  return guarded(ctx, [&] {
    G5_CHECK(t, "g5_batch_get_tuning: null output");
    t->alpha = ctx->ms.alpha;
    t->pull_grid = ctx->ms.pull_grid;
    t->sweep_below = ctx->ms.sweep_below;
    t->sweep_grid = ctx->ms.sweep_grid;
    t->persistent = ctx->ms.persistent;
    t->persist_levels = ctx->ms.persist_levels;
    t->persist_grid = ctx->ms.persist_grid;
    t->persist_waves = ctx->ms.persist_waves;
    t->pull_filter = ctx->ms.pull_filter;
    t->beta = ctx->ms.beta;
    t->compact_below = ctx->ms.compact_below;
    t->filter_from = ctx->ms.filter_from;
    t->pull_unroll = ctx->ms.pull_unroll;
    t->lane_rule = ctx->ms.lane_rule;
    t->mixed_push = ctx->ms.mixed_push;
    t->reserved = 0;
  });
}

int g5_run_bfs_batch(g5_ctx* ctx, const int64_t* roots, int num_roots) {
  return guarded(ctx, [&] {
    G5_CHECK(roots, "g5_run_bfs_batch: null roots");
    set_device(ctx);
    ms_run(ctx, roots, num_roots);
  });
}

int g5_batch_get_levels(g5_ctx* ctx, int lane, uint8_t* levels) {
  return guarded(ctx, [&] {
    G5_CHECK(levels, "g5_batch_get_levels: null output");
    set_device(ctx);
    ms_get_levels(ctx, lane, levels);
  });
}

int g5_batch_teps_edges(g5_ctx* ctx, int64_t* counts) {
  return guarded(ctx, [&] {
    G5_CHECK(counts, "g5_batch_teps_edges: null output");
    set_device(ctx);
    ms_teps_edges(ctx, counts);
  });
}

int g5_batch_stats_get(g5_ctx* ctx, g5_batch_stats* out) {
  return guarded(ctx, [&] {
    G5_CHECK(out, "g5_batch_stats_get: null output");
    if (!ctx->ms.have_result) throw Error{"g5_batch_stats_get: no batch result", G5_ERR_STATE};
    *out = ctx->ms.stats;
  });
}

int g5_batch_model_bytes(g5_ctx* ctx, double* per_level, double* total) {
  return guarded(ctx, [&] {
    set_device(ctx);
    ms_model_bytes(ctx, per_level, total);
  });
}

int g5_batch_level_lanes(g5_ctx* ctx, uint64_t* push_lanes, uint64_t* pull_lanes) {
  return guarded(ctx, [&] {
    const MsState& m = ctx->ms;
    G5_CHECK(m.have_result, "g5_batch_level_lanes: no batch result");
    G5_CHECK(push_lanes && pull_lanes, "g5_batch_level_lanes: null output");
    const uint64_t all = m.nroots == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << m.nroots) - 1);
    for (int l = 0; l < G5_MAX_LEVELS; ++l) {
      const bool run = l >= 1 && l < m.stats.num_levels;
      push_lanes[l] = run && m.level_masks ? m.level_push[l] & all : 0;
      pull_lanes[l] = run ? (m.level_masks ? m.level_pull[l] & all : all) : 0;
    }
  });
}

int g5_batch_end(g5_ctx* ctx) {
  return guarded(ctx, [&] {
    set_device(ctx);
    G5_HIP(hipStreamSynchronize(ctx->stream));
    ctx->ms.release();
  });
}

}  // extern "C"
