// g5_ctx.hpp -- the per-GPU context behind the C ABI.
#pragma once

#include "g5_internal.hpp"
#include "g5_publish.hpp"

#include <memory>

namespace g5 {

// Canonical CSR of one graph in HBM (see g5_internal.hpp for the layout).
struct Graph {
  bool built = false;
  int scale = 0;             // estimated SCALE (msb of OR of all ids + 1)
  int reorder_mode = 2;
  int64_t n = 0;             // 2^scale
  int64_t n_active = 0;      // vertices with >= 1 non-loop tuple
  int64_t n_rows = 0;        // rows swept by the BFS: n_active (modes 1,2) or n (mode 0)
  int64_t num_tuples = 0;
  int64_t self_loops = 0;
  int64_t directed_before = 0;
  int64_t nnz = 0;
  int64_t max_degree = 0;
  int64_t max_used_vertex = 0;
  double construct_ms = 0;
  DevBuf<uint64_t> offsets;    // n_rows + 1
  DevBuf<uint32_t> cols;       // nnz
  DevBuf<uint32_t> first_nbr;  // n_rows; 0xFFFFFFFF for an empty row
  DevBuf<uint32_t> weight;     // n_rows: non-loop degree + 2 * self loops
  DevBuf<uint32_t> old2new;    // n
  DevBuf<uint32_t> new2old;    // n
  std::vector<uint64_t> has_edge_words;  // host bitmap by original id
  void clear() {
    built = false;
    offsets.release();
    cols.release();
    first_nbr.release();
    weight.release();
    old2new.release();
    new2old.release();
    has_edge_words.clear();
  }
};

// Single-root BFS state.
struct BfsState {
  bool prepared = false;
  bool have_result = false;
  int64_t root_orig = -1;
  uint32_t root_new = 0;
  int64_t words = 0;              // bitmap words covering n_rows
  DevBuf<uint64_t> vis[2];
  DevBuf<uint8_t> level;          // n_rows
  DevBuf<uint32_t> parent;        // n_rows
  DevBuf<uint32_t> queue[2];      // n_rows
  DevBuf<uint64_t> prefix;        // n_rows + 1 (top-down edge prefix)
  DevBuf<uint32_t> chunk_rows;    // 2 per top-down edge chunk: first/last frontier index
  DevBuf<uint64_t> scan_tmp;
  DevBuf<uint64_t> counters;      // device counters (see g5_bfs.hip)
  uint64_t* h_counters = nullptr; // pinned host mirror
  PublishBlock pub;               // the level counters, published by k_publish
  uint64_t pub_seq = 0;
  PublishBlock fuse;              // the fused small levels' state (g5_bfs.hip)
  uint64_t fuse_seq = 0;
  int64_t small_edges = 8192;     // small top-down levels up to this many edges run fused (0: never)
  int64_t tail_pull = 4;          // post-bottom-up top-down levels with tail_pull * nf > unvisited run bottom-up
  DevBuf<int64_t> pred_out;       // n, built on demand by get_pred
  DevBuf<uint8_t> level_out;      // n, built on demand by get_levels
  int64_t bin_edges = 0;  // top-down levels with at least this many edges are binned (0: never)
  DevBuf<uint64_t> bin_pairs;     // (target << 32 | source) of a binned level
  DevBuf<uint64_t> bin_cnt;       // bucket counts and cursors
  int64_t td_pieces = 0;          // pieces per row of a piece-ordered hub level (0: row order)
  DevBuf<uint64_t> seg_pre;       // segment list of a piece-ordered level (g5_bfs.hip)
  DevBuf<uint64_t> seg_base;
  DevBuf<uint32_t> seg_src;
  DevBuf<uint64_t> seg_count;
  DevBuf<uint64_t> seg_scan_tmp;
  DevBuf<uint64_t> seg_start;     // first cols index of each row suffix
  int64_t td_head = 4;            // head threshold factor in 1/16 (0: no head; see k_td_head_size)
  int64_t seg_max = 0;            // segments per segment-list level (0: the count field's limit)
  int64_t hub_shift = 0;          // target block = 2^hub_shift ids (0: no block index)
  int64_t hub_built_shift = 0;    // shift the index was built for (0: none)
  int64_t hub_pieces = 0;         // blocks + 1
  int64_t hub_rows = 0;           // rows [0, hub_rows) are indexed
  DevBuf<uint32_t> hub_bnd;       // hub_rows * hub_pieces
  g5_bfs_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> lev_begin, lev_end;  // per level
  // the tuning (g5_bfs_set_tuning's fields) of another state
  void copy_tuning(const BfsState& o) {
    small_edges = o.small_edges;
    tail_pull = o.tail_pull;
    bin_edges = o.bin_edges;
    td_pieces = o.td_pieces;
    td_head = o.td_head;
    seg_max = o.seg_max;
    hub_shift = o.hub_shift;
  }
  void release() {
    for (auto& v : vis) v.release();
    level.release();
    parent.release();
    for (auto& q : queue) q.release();
    prefix.release();
    chunk_rows.release();
    scan_tmp.release();
    counters.release();
    pred_out.release();
    level_out.release();
    bin_pairs.release();
    bin_cnt.release();
    seg_pre.release();
    seg_base.release();
    seg_src.release();
    seg_count.release();
    seg_scan_tmp.release();
    seg_start.release();
    hub_bnd.release();
    hub_built_shift = 0;
    if (h_counters) (void)hipHostFree(h_counters);
    h_counters = nullptr;
    pub.release();
    fuse.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    ev0 = ev1 = nullptr;
    for (auto e : lev_begin) (void)hipEventDestroy(e);
    for (auto e : lev_end) (void)hipEventDestroy(e);
    lev_begin.clear();
    lev_end.clear();
    prepared = false;
    have_result = false;
  }
};

// Batched (up to 64 roots) BFS state.
struct MsState {
  bool prepared = false;
  bool have_result = false;
  int nroots = 0;
  int nlevels = 0;
  std::vector<int64_t> host_roots;
  double alpha = 128.0;      // direction threshold (its meaning depends on lane_rule)
  int64_t pull_grid = 8192;  // workgroups of the vertex-parallel pull
  double sweep_below = 0.0;  // edge-parallel sweep while reached pairs < this x roots x rows
  int64_t sweep_grid = 4096; // workgroups of the edge-parallel sweep
  uint64_t sweep_chunks = 0;
  DevBuf<uint32_t> sweep_rows;  // first/last row of every sweep chunk of the CSR
  DevBuf<uint32_t> chunk_list;  // sweep chunks holding a frontier row
  // persistent level-fused path (one hipGraph replay per batch)
  int persistent = 0;
  int persist_levels = 16;     // frontier-word arrays preallocated for it
  int64_t persist_grid = 0;    // 0: every co-resident workgroup
  int persist_waves = 4;       // register budget of the persistent kernel: 4, 6 or 8 waves per SIMD
  int pull_filter = 0;         // pull tests the nonzero bitmap before a frontier word: 0 never, 1 always,
                               // 2 for vertices from filter_from x rows on
  double filter_from = 0.01;
  int pull_unroll = 0;         // probe unrolling of the per-level pulls (0..7, see g5_msbfs.hip)
  int lane_rule = 1;           // 0: one direction per level for every root; 1: per root (per-level launches)
  int mixed_push = 1;          // push of a mixed level: 0 over CSR chunks, 1 over a queue of the frontier vertices
  uint64_t level_push[G5_MAX_LEVELS] = {};  // per level: roots pushed, roots pulled (per-root rule)
  uint64_t level_pull[G5_MAX_LEVELS] = {};
  bool level_masks = false;    // the last run used the per-root rule
  double beta = 50.0;          // pull whenever fewer than 1/beta of the (root, vertex) pairs are unreached
  double compact_below = 0.0;  // pull from a compact frontier when it spans fewer than this x rows vertices
  DevBuf<uint64_t> pack, compact;
  DevBuf<uint64_t> nz_sel;     // per-root rule: vertices holding a pushing root's frontier
  hipGraphExec_t graph_exec = nullptr;
  int graph_levels = 0;
  double graph_alpha = 0, graph_sweep = 0;
  int64_t graph_grid = 0;
  int graph_waves = 0, graph_filter = -1;
  double graph_beta = -1;
  unsigned p_grid = 0;
  int clock_khz = 100000;
  DevBuf<uint64_t*> p_fr;
  DevBuf<uint64_t> p_cnt, p_misc, p_bar;
  DevBuf<int64_t> p_ts;
  DevBuf<int> p_dir;
  uint64_t* h_persist = nullptr;
  alignas(16) unsigned char p_args_storage[256];  // kernel argument block of the persistent launch
  DevBuf<uint64_t> seen;     // n_rows
  DevBuf<uint64_t> nz[2];    // frontier-word-not-zero bitmaps, alternating by level
  std::vector<std::unique_ptr<DevBuf<uint64_t>>> fr;  // frontier words per level
  // per level: one bit per vertex whose frontier word the level's last run
  // left non-zero (a copy of that level's nonzero bitmap); valid where
  // fr_exact is set, so the next batch clears only those words
  std::vector<std::unique_ptr<DevBuf<uint64_t>>> fr_dirty;
  std::vector<char> fr_exact;
  std::vector<int64_t> fr_count;  // per level: vertices the last run marked
  DevBuf<const uint64_t*> frp;  // device copy of the fr pointers
  DevBuf<uint32_t> queue;
  DevBuf<uint32_t> touch;    // vertices a sparse push level touched
  DevBuf<uint64_t> miss;     // one bit per vertex still missing a live root (tail pull levels)
  DevBuf<uint32_t> ulist;    // the same vertices as a list
  DevBuf<uint64_t> ulist_count;
  DevBuf<uint64_t> prefix;
  DevBuf<uint32_t> chunk_rows;
  DevBuf<uint64_t> scan_tmp;
  DevBuf<uint64_t> counters;
  DevBuf<int64_t> roots;
  uint64_t* h_counters = nullptr;
  PublishBlock pub;             // the counters as each level publishes them (g5_msbfs.hip)
  uint64_t pub_seq = 0;
  g5_batch_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> lev_begin, lev_end;
  void release() {
    seen.release();
    for (auto& b : nz) b.release();
    fr.clear();
    fr_dirty.clear();
    fr_exact.clear();
    fr_count.clear();
    frp.release();
    queue.release();
    touch.release();
    miss.release();
    ulist.release();
    ulist_count.release();
    prefix.release();
    chunk_rows.release();
    scan_tmp.release();
    counters.release();
    roots.release();
    sweep_rows.release();
    sweep_chunks = 0;
    chunk_list.release();
    pack.release();
    compact.release();
    nz_sel.release();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    graph_levels = 0;
    p_fr.release();
    p_cnt.release();
    p_misc.release();
    p_bar.release();
    p_ts.release();
    p_dir.release();
    if (h_persist) (void)hipHostFree(h_persist);
    h_persist = nullptr;
    if (h_counters) (void)hipHostFree(h_counters);
    h_counters = nullptr;
    pub.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    ev0 = ev1 = nullptr;
    for (auto e : lev_begin) (void)hipEventDestroy(e);
    for (auto e : lev_end) (void)hipEventDestroy(e);
    lev_begin.clear();
    lev_end.clear();
    prepared = false;
    have_result = false;
  }
};

}  // namespace g5

struct g5_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // device edge list
  g5::DevBuf<g5_packed_edge> edges;
  int64_t num_edges = 0;
  g5::DevBuf<g5_weighted_edge> wstage;  // weighted tuples on their way to the host
  double generate_ms = 0;
  g5::Graph graph;
  g5::BfsState bfs;
  g5::MsState ms;
  g5::DevBuf<uint64_t> acc;  // small reduction scratch
  int servers = 0;           // g5_server objects that serve this context's graph
};

namespace g5 {
// g5_csr.hip
void build_graph(g5_ctx* ctx, int reorder_mode, bool release_edges);
// g5_gen.hip
void launch_initial_tree(g5_packed_edge* out, int64_t count, int scale, uint64_t s1, uint64_t s2,
                         hipStream_t stream);
void launch_kronecker(int device, g5_packed_edge* out, int64_t start, int64_t count, int scale,
                      uint64_t s1, uint64_t s2, int a_num, int bc_num, hipStream_t stream);
void launch_weighted(int device, const g5_packed_edge* in, int64_t start_index, int64_t count, int max_weight,
                     uint64_t s1, uint64_t s2, g5_weighted_edge* out, hipStream_t stream);
// g5_bfs.hip
void bfs_prepare(g5_ctx* ctx);
void bfs_tuning_buffers(g5_ctx* ctx);
void bfs_run(g5_ctx* ctx, int64_t root_orig);
void bfs_get_pred(g5_ctx* ctx, int64_t* host_pred);
void bfs_get_levels(g5_ctx* ctx, uint8_t* host_levels);
int64_t bfs_teps_edges(g5_ctx* ctx);
// the same on a BFS state and stream other than the context's own (the
// graph is the context's; several states may run on one graph at once)
void bfs_prepare(g5_ctx* ctx, BfsState& b, hipStream_t s);
void bfs_tuning_buffers(g5_ctx* ctx, BfsState& b, hipStream_t s);
void bfs_run(g5_ctx* ctx, BfsState& b, hipStream_t s, int64_t root_orig);
int64_t bfs_teps_edges(g5_ctx* ctx, const BfsState& b, DevBuf<uint64_t>& acc, hipStream_t s);
void bfs_model_bytes(g5_ctx* ctx, double* per_level, double* total);
// g5_validate.hip
void bfs_validate(g5_ctx* ctx, int64_t* errors_out);
int64_t pred_validate(g5_ctx* ctx, int64_t root, const int64_t* host_pred, int64_t* errors_out);
int64_t pred_validate_noalloc(g5_ctx* ctx, int64_t root, const int64_t* host_pred, int64_t* errors_out);
// g5_msbfs.hip
void ms_prepare(g5_ctx* ctx);
void ms_alloc_compact(g5_ctx* ctx);
void ms_run(g5_ctx* ctx, const int64_t* roots, int nroots);
void ms_get_levels(g5_ctx* ctx, int lane, uint8_t* host_levels);
void ms_teps_edges(g5_ctx* ctx, int64_t* out);
void ms_model_bytes(g5_ctx* ctx, double* per_level, double* total);
}  // namespace g5
