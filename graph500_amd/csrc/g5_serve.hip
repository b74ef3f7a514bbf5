// g5_serve.hip -- serving loop: (graph, root) BFS queries over graphs resident
// in HBM, with continuous batching.
//
// Each graph has an engine: the 64-lane bit-packed BFS of the batched path
// (one lane per query), run one level at a time on the graph's own stream, so
// the levels of different graphs overlap on the device.  At every level
// boundary the engine
//   * retires the lanes that discovered nothing in the level just finished
//     (their answers are complete; their traversed-edge counts are taken from
//     the seen words before the lane is reused),
//   * admits waiting queries into free lanes while answer memory allows (one
//     n_rows-byte level array per query, from a preallocated pool), placing
//     each root into the current frontier at the current level,
//   * launches the next level for every active lane, whatever level each
//     lane's own search has reached.
// A lane's level values are written into its query's answer array as they
// are discovered (level = engine level - the level the lane was admitted
// at), so an answer needs no extraction.  A retired lane "cools" for one
// level: the next level's sweep clears its bits from the seen words.
//
// Reference: there is no serving loop in the reference; each answer equals
// BfsBase::run_bfs's levels for the same root (mpi/bfs.hpp:2651-2940).
#include "g5_ctx.hpp"
#include "g5_ms_device.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <string>
#include <thread>
#include <vector>

namespace g5 {

// Epilogue of a serving level: writes each newly reached (lane, vertex) into
// the lane's answer array (a wave writes 64 consecutive bytes per lane),
// counts each lane's new vertices and sums their row lengths (lane_nq[0, 64)
// and [64, 128)), and marks the vertices with new lanes among `sel` (the
// top-down lanes) in nz_sel.  The parameters stay in the kernel argument
// block; ServeLocal holds the per-thread state.
struct ServeEpi {
  uint8_t* ans[64];
  uint8_t lvl[64];
  unsigned long long* lane_nq;
  const uint64_t* offsets;
  uint64_t* nz_sel;
  uint64_t sel;
  uint64_t stat_lanes;  // lanes whose counts and degree sums the host still needs
  int list;             // the fold walks a list of vertices, not every row's word
};

// Thread i of a wave keeps lane i's answer pointer and level in registers
// (read by readlane where a wave writes lane r's bytes) and lane i's counts
// (a readlane walk over the wave's vertices, as LaneEpi), so no shared
// counter is touched until the block's LDS reduction at the end.
struct ServeLocal {
  const ServeEpi& p;
  mutable uint64_t n_acc = 0, d_acc = 0, my_ans = 0;
  mutable uint32_t my_lvl = 0;
  __device__ unsigned long long* lds() const {
    __shared__ unsigned long long s_nq[128];
    return s_nq;
  }
  __device__ void begin() const {
    const int me = threadIdx.x & 63;
    my_ans = (uint64_t)p.ans[me];
    my_lvl = p.lvl[me];
    if (threadIdx.x < 128) lds()[threadIdx.x] = 0;
    __syncthreads();
  }
  __device__ void operator()(int64_t v, bool inb, uint64_t nw) const {
    const int me = threadIdx.x & 63;
    uint64_t u = wave_or(inb ? nw : 0);
    while (u) {
      const int r = __builtin_ctzll(u);
      u &= u - 1;
      const uint32_t alo = __builtin_amdgcn_readlane((uint32_t)my_ans, r);
      const uint32_t ahi = __builtin_amdgcn_readlane((uint32_t)(my_ans >> 32), r);
      const uint32_t lv = __builtin_amdgcn_readlane(my_lvl, r);
      if (inb && ((nw >> r) & 1)) ((uint8_t*)(((uint64_t)ahi << 32) | alo))[v] = (uint8_t)lv;
    }
    const uint64_t want = inb ? nw & p.stat_lanes : 0;
    const uint64_t deg = want ? p.offsets[v + 1] - p.offsets[v] : 0;
    uint64_t m = __ballot(want != 0);
    while (m) {
      const int t = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const uint32_t wlo = __builtin_amdgcn_readlane((uint32_t)want, t);
      const uint32_t whi = __builtin_amdgcn_readlane((uint32_t)(want >> 32), t);
      const uint32_t dlo = __builtin_amdgcn_readlane((uint32_t)deg, t);
      const uint32_t dhi = __builtin_amdgcn_readlane((uint32_t)(deg >> 32), t);
      const uint32_t bit = me < 32 ? (wlo >> me) & 1 : (whi >> (me - 32)) & 1;
      if (bit) {
        ++n_acc;
        d_acc += ((uint64_t)dhi << 32) | dlo;
      }
    }
    if (p.list) {
      // a listed fold: nz_sel was cleared, its bits are set one vertex at a time
      if (inb && (nw & p.sel) != 0)
        atomicOr((unsigned long long*)(p.nz_sel + (v >> 6)), 1ull << (v & 63));
    } else {
      const uint64_t bs = __ballot(inb && (nw & p.sel) != 0);
      if (me == 0) p.nz_sel[v >> 6] = bs;
    }
  }
  __device__ void end() const {
    const int me = threadIdx.x & 63;
    if (n_acc) {
      atomicAdd(&lds()[me], (unsigned long long)n_acc);
      atomicAdd(&lds()[64 + me], (unsigned long long)d_acc);
    }
    __syncthreads();
    if (threadIdx.x < 128 && lds()[threadIdx.x]) atomicAdd(p.lane_nq + threadIdx.x, lds()[threadIdx.x]);
  }
};

// A level's counters into the engine's coherent host block ([0, M_NUM) the
// level counters, [M_NUM, M_NUM + 128) the per-lane counts and degree sums,
// [kSvTeps, kSvTeps + 64) the per-lane weight sums of the last retirement),
// each tagged with the level's sequence number: the host learns that the
// level finished, and its counters, by polling host memory, with no copy
// commands on the stream.
constexpr int kSvTeps = M_NUM + 128;
constexpr int kSvPub = kSvTeps + 64;
constexpr int kSvTouch = M_NUM + 1 + 128;  // device counter slot of a listed push's touched vertices
constexpr int kSvCounters = kSvTouch + 1;
__global__ void k_sv_publish(const unsigned long long* __restrict__ c, const unsigned long long* __restrict__ tacc,
                             Published* __restrict__ p, uint64_t seq) {
  const int t = threadIdx.x;
  if (t < M_NUM) publish(p + t, c[t], seq);
  if (t < 128) publish(p + M_NUM + t, c[M_NUM + 1 + t], seq);
  if (t < 64) publish(p + kSvTeps + t, tacc[t], seq);
}

struct ServeRoots {
  int64_t root[64];  // original id of the query admitted into each lane, -1 = none
  uint8_t* ans[64];
};

__global__ void __launch_bounds__(kBlock) k_sv_pull(uint64_t* __restrict__ seen, const uint64_t* __restrict__ front,
                                                    const uint64_t* __restrict__ nz_front, uint64_t* __restrict__ next,
                                                    uint64_t* __restrict__ nz_next,
                                                    const uint32_t* __restrict__ first_nbr,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ cols, int64_t rows, uint64_t live_in,
                                                    int filter, unsigned long long* __restrict__ c, uint64_t keep,
                                                    int merge, ServeEpi epi) {
  const FrontView fv{front, nz_front, nullptr, nullptr, filter ? kFrontFilter : kFrontPlain, 0};
  const ServeLocal el{epi};
  // two probes per step on a short row, one per lane and step on a long row:
  // the batch's fastest pull unrolling at SCALE 28 (fewer probes past the
  // point where a vertex's searches are all found)
  ms_pull_body<ServeLocal, 2, 1>(seen, fv, next, nz_next, first_nbr, offsets, cols, rows, live_in, c, blockIdx.x,
                                 gridDim.x, keep, el, merge != 0);
}

__global__ void __launch_bounds__(kBlock) k_sv_finish(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                                      uint64_t* __restrict__ nz_next,
                                                      const uint64_t* __restrict__ offsets, int64_t rows,
                                                      unsigned long long* __restrict__ c, uint64_t keep,
                                                      ServeEpi epi) {
  const ServeLocal el{epi};
  ms_finish_body(seen, next, nz_next, offsets, rows, c, blockIdx.x, gridDim.x, keep, el);
}

__global__ void __launch_bounds__(kBlock) k_sv_collect(const uint32_t* __restrict__ chunk_rows, uint64_t nchunks,
                                                       const uint64_t* __restrict__ nz_front,
                                                       uint32_t* __restrict__ list, unsigned* count) {
  ms_collect_chunks_body(chunk_rows, nchunks, nz_front, list, count, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_sv_push(const uint32_t* __restrict__ list, const unsigned* count,
                                                    const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ cols,
                                                    const uint64_t* __restrict__ front,
                                                    const uint64_t* __restrict__ nz_front, uint64_t* next,
                                                    uint64_t lanes) {
  __shared__ SweepLds L;
  ms_push_list_body(list, count, chunk_rows, nnz, offsets, cols, front, nz_front, next, L, blockIdx.x, gridDim.x,
                    lanes);
}

__global__ void __launch_bounds__(kBlock) k_sv_push_touch(const uint32_t* __restrict__ list, const unsigned* count,
                                                          const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                                          const uint64_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ cols,
                                                          const uint64_t* __restrict__ front,
                                                          const uint64_t* __restrict__ nz_front, uint64_t* next,
                                                          uint64_t lanes, uint32_t* __restrict__ touch,
                                                          unsigned long long* __restrict__ touch_count) {
  __shared__ SweepLds L;
  const unsigned n = __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (unsigned i = blockIdx.x; i < n; i += gridDim.x)
    push_chunk_touch(list[i], chunk_rows, nnz, offsets, cols, front, nz_front, next, L, lanes, touch, touch_count);
}

// The fold of a small push level over the vertices it touched (no lane
// cooling: the full fold also clears cooling lanes from seen).
__global__ void __launch_bounds__(kBlock) k_sv_finish_list(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                                           uint64_t* __restrict__ nz_next,
                                                           const uint64_t* __restrict__ offsets,
                                                           const uint32_t* __restrict__ touch,
                                                           const unsigned long long* __restrict__ touch_count,
                                                           unsigned long long* __restrict__ c, ServeEpi epi) {
  const ServeLocal el{epi};
  ms_finish_list_body(seen, next, nz_next, offsets, touch, (uint64_t)*touch_count, c, blockIdx.x, gridDim.x, el);
}

// Zeroes the frontier words marked in a nonzero bitmap (one bitmap word per
// lane; for sparse marks).
__global__ void __launch_bounds__(kBlock) k_sv_clear_marked(const uint64_t* __restrict__ marked, int64_t words,
                                                            int64_t rows, uint64_t* __restrict__ fr) {
  for (int64_t wd = (int64_t)blockIdx.x * kBlock + threadIdx.x; wd < words; wd += (int64_t)gridDim.x * kBlock)
    for (uint64_t b = marked[wd]; b; b &= b - 1) {
      const int64_t v = wd * 64 + __builtin_ctzll(b);
      if (v < rows) fr[v] = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_sv_select(const uint64_t* __restrict__ front, int64_t rows, uint64_t lanes,
                                                      uint64_t* __restrict__ nz_sel) {
  ms_select_body(front, rows, lanes, nz_sel, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_sv_zero(uint64_t* __restrict__ a, int64_t n) {
  ms_zero_body(a, n, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_sv_fill_bytes(uint8_t* __restrict__ p, int64_t n, uint8_t val) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = val;
}

// seen &= keep over all rows (lanes cooling while the engine is idle).
__global__ void __launch_bounds__(kBlock) k_sv_clear(uint64_t* __restrict__ seen, int64_t rows, uint64_t keep) {
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += (int64_t)gridDim.x * kBlock) {
    const uint64_t s = seen[v];
    if (s & ~keep) seen[v] = s & keep;
  }
}

// Places the admitted roots into seen / frontier / nonzero bitmap and their
// answers; one thread per lane (lanes may share a root vertex).
__global__ void k_sv_admit(const uint32_t* __restrict__ old2new, int64_t rows, ServeRoots rs,
                           uint64_t* __restrict__ seen, uint64_t* __restrict__ front, uint64_t* __restrict__ nz_front,
                           uint64_t* __restrict__ nz_sel_front) {
  const int r = threadIdx.x;
  if (r >= 64 || rs.root[r] < 0) return;
  const uint32_t v = old2new[rs.root[r]];
  if ((int64_t)v >= rows) return;  // a root without edges reaches nothing
  const unsigned long long bit = 1ull << r;
  atomicOr((unsigned long long*)(seen + v), bit);
  atomicOr((unsigned long long*)(front + v), bit);
  atomicOr((unsigned long long*)(nz_front + (v >> 6)), 1ull << (v & 63));
  atomicOr((unsigned long long*)(nz_sel_front + (v >> 6)), 1ull << (v & 63));
  rs.ans[r][v] = 0;
}

// Per-lane sums of the tuple weights of the reached vertices, lanes in mask.
__global__ void __launch_bounds__(kBlock) k_sv_teps(const uint64_t* __restrict__ seen,
                                                    const uint32_t* __restrict__ weight, int64_t rows, uint64_t mask,
                                                    unsigned long long* __restrict__ acc) {
  // vertices reached by every lane of mask add to one sum; for the others,
  // thread i of a wave sums lane i's weights (readlane walk over the wave's
  // vertices, no shared counters until the block reduction)
  __shared__ unsigned long long s_acc[65];
  if (threadIdx.x < 65) s_acc[threadIdx.x] = 0;
  __syncthreads();
  const int me = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
  uint64_t full = 0, mine = 0;
  for (int64_t wd = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); wd * 64 < rows; wd += nwaves) {
    const int64_t v = wd * 64 + me;
    const uint64_t x = v < rows ? seen[v] & mask : 0;
    const uint64_t w = x ? weight[v] : 0;
    if (x == mask) full += w;
    uint64_t m = __ballot(x != 0 && x != mask);
    while (m) {
      const int t = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const uint32_t xlo = __builtin_amdgcn_readlane((uint32_t)x, t);
      const uint32_t xhi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), t);
      const uint32_t wt = __builtin_amdgcn_readlane((uint32_t)w, t);
      const uint32_t bit = me < 32 ? (xlo >> me) & 1 : (xhi >> (me - 32)) & 1;
      if (bit) mine += wt;
    }
  }
  full = wave_sum(full);
  if (me == 0 && full) atomicAdd(&s_acc[64], (unsigned long long)full);
  if (mine) atomicAdd(&s_acc[me], (unsigned long long)mine);
  __syncthreads();
  if (threadIdx.x < 64) {
    const unsigned long long t = s_acc[threadIdx.x] + (((mask >> threadIdx.x) & 1) ? s_acc[64] : 0);
    if (t) atomicAdd(acc + threadIdx.x, t);
  }
}

// Answer (new ids) -> levels by original id.
__global__ void k_sv_levels(const uint32_t* __restrict__ old2new, int64_t n, int64_t rows,
                            const uint8_t* __restrict__ ans, int64_t root, uint8_t* __restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t nv = old2new[v];
    out[v] = (v == root) ? 0 : ((int64_t)nv < rows ? ans[nv] : 0xFF);
  }
}

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum { SV_WAITING = 0, SV_RUNNING = 1, SV_DONE = 2, SV_RELEASED = 3 };

struct Request {
  int graph = 0;
  int64_t root = 0;
  double arrival = 0, admit = -1, done = -1;
  int lane = -1;
  int buf = -1;
  int64_t edges = -1;
  int status = SV_WAITING;
  bool keep = false;
  int levels = 0;  // levels the query's lane took part in
  int64_t result_slot = -1;
};

// A slot that answers one query at a time with the single-root kernels
// (config solo), driven by a host thread of its own so that the server loop
// goes on with the other graphs and slots; the thread owns the slot from
// state 1 until it sets state 2.  Slot 0 runs on the graph context's own BFS
// state and stream, every further slot (config solo_slots) on a state and a
// stream of its own, so that the slots of one graph overlap on the device.
struct SoloSlot {
  BfsState own;              // the state of a slot beyond the first
  BfsState* b = nullptr;     // the state the runs use
  hipStream_t s = nullptr;
  bool own_stream = false;
  DevBuf<uint64_t> acc;      // the traversed-edge reduction's scratch
  std::thread th;
  std::atomic<int> state{0};  // 0 free, 1 running, 2 finished
  int64_t req = -1;
  int64_t edges = 0;
  int levels = 0;
  double done = 0;  // seconds after the run's start
  std::string err;

  ~SoloSlot() {
    if (th.joinable()) th.join();
    own.release();
    if (own_stream && s) (void)hipStreamDestroy(s);
  }
};

struct Engine {
  g5_ctx* ctx = nullptr;
  hipStream_t s = nullptr;
  int64_t n = 0, rows = 0, words = 0, nnz = 0;
  uint64_t nchunks = 0;
  DevBuf<uint64_t> seen, fr[2], nz[2];
  DevBuf<uint64_t> nzs[2];  // nonzero bitmap restricted to the lanes that were top-down
  uint64_t nzs_lanes[2] = {0, 0};  // the lanes whose frontier vertices nzs[i] marks
  DevBuf<uint32_t> chunk_rows, chunk_list;
  DevBuf<unsigned long long> cnt;    // M_NUM counters, chunk count, 64 + 64 per-lane counts and degree
                                     // sums, the touched-vertex count of a listed push (kSvTouch)
  DevBuf<uint32_t> touch;            // rows: the vertices a listed push level touched
  int64_t nzc[2] = {0, 0};           // vertices marked in nz[i] (the frontier words non-zero in fr[i])
  DevBuf<unsigned long long> tacc;   // per-lane tuple-weight sums of a retirement
  unsigned long long* h = nullptr;   // pinned: the per-lane weight sums of a retirement, copied at the end
  PublishBlock pub;                  // the level counters and per-lane counts (k_sv_publish)
  uint64_t pub_seq = 0;              // sequence number of the level in flight
  hipEvent_t ev_start = nullptr, ev_end = nullptr;  // timing of the level in flight
  g5_server_graph_stats st{};
  std::vector<std::unique_ptr<DevBuf<uint8_t>>> bufs;
  std::vector<int> free_bufs;
  int64_t lane_req[64];
  int lane_l0[64];
  // per-lane direction state, the reference rule per search (mpi/bfs.hpp:2799-2829)
  bool lane_fwd[64];
  uint64_t pulled = 0;  // lane rule 3: searches that have pulled keep pulling
  bool lane_growing[64];
  int64_t lane_prev_nq[64];
  int lane_pulls[64];    // pull levels each search has taken
  int64_t lane_visited[64];
  double lane_vdeg[64];  // row lengths of the vertices each search has reached
  uint64_t active = 0, cooling = 0;
  int level = 0;        // fr[level & 1] is the current frontier
  bool busy = false;    // a level is in flight
  bool teps_pending = false;
  uint64_t teps_mask = 0;
  int64_t teps_req[64];
  uint64_t last_fdeg = 0;
  std::deque<int64_t> waiting;
  int64_t levels_run = 0;
  int64_t admit_events = 0;
  // single-root run slots (config solo, solo_slots)
  std::vector<std::unique_ptr<SoloSlot>> solo;

  bool solo_running() const {
    for (auto& sl : solo)
      if (sl->state.load(std::memory_order_acquire) != 0) return true;
    return false;
  }

  ~Engine() {
    solo.clear();  // joins the runs still going
    if (h) (void)hipHostFree(h);
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_end) (void)hipEventDestroy(ev_end);
  }
};

}  // namespace

}  // namespace g5

struct g5_server {
  std::string err;
  std::vector<std::unique_ptr<g5::Engine>> eng;
  std::vector<g5::Request> req;
  g5_server_config cfg{};
  int device = 0;
  int64_t done_count = 0;
  int64_t run_base = 0;  // first request of the current run
  int64_t run_done = 0;  // requests of the current run answered so far
  double t0 = 0;  // start of the current run (now_s clock)
  g5::DevBuf<uint8_t> lv_out;
};

namespace g5 {

// The single-root slots of a graph (config solo): slot 0 on the context's own
// BFS state and stream, the others on states of their own with the context's
// BFS tuning.
static void sv_setup_solo(Engine& e, int slots) {
  g5_ctx* ctx = e.ctx;
  if (!ctx->bfs.prepared) bfs_prepare(ctx);
  for (int i = 0; i < slots; ++i) {
    std::unique_ptr<SoloSlot> sl(new SoloSlot());
    if (i == 0) {
      sl->b = &ctx->bfs;
      sl->s = ctx->stream;
    } else {
      sl->own.copy_tuning(ctx->bfs);
      G5_HIP(hipStreamCreate(&sl->s));
      sl->own_stream = true;
      bfs_prepare(ctx, sl->own, sl->s);
      sl->b = &sl->own;
    }
    sl->acc.alloc(4);
    e.solo.push_back(std::move(sl));
  }
}

static void sv_setup_engine(g5_server* srv, Engine& e, g5_ctx* ctx) {
  Graph& g = ctx->graph;
  G5_CHECK(g.built, "g5_server_create: a graph context holds no graph (call g5_construct first)");
  G5_CHECK(g.reorder_mode != 0, "g5_server_create: graphs must be built with reorder mode 1 or 2");
  e.ctx = ctx;
  e.s = ctx->stream;
  e.n = g.n;
  e.rows = g.n_rows;
  e.words = (e.rows + 63) / 64;
  e.nnz = g.nnz;
  e.seen.alloc(e.rows);
  for (auto& b : e.fr) b.alloc(e.rows);
  for (auto& b : e.nz) b.alloc(e.words);
  for (auto& b : e.nzs) b.alloc(e.words);
  e.nchunks = ((uint64_t)e.nnz + kSweepEpb - 1) / kSweepEpb;
  e.chunk_rows.alloc(2 * e.nchunks + 2);
  e.chunk_list.alloc(e.nchunks + 1);
  e.cnt.alloc(kSvCounters);
  e.touch.alloc(e.rows ? e.rows : 1);
  e.tacc.alloc(64);
  G5_HIP(hipMemsetAsync(e.tacc.p, 0, 64 * 8, e.s));
  G5_HIP(hipHostMalloc((void**)&e.h, 64 * sizeof(unsigned long long), hipHostMallocDefault));
  e.pub.alloc(kSvPub);
  G5_HIP(hipEventCreate(&e.ev_start));
  G5_HIP(hipEventCreate(&e.ev_end));
  hipStream_t s = e.s;
  if (e.nchunks)
    hipLaunchKernelGGL(k_chunk_rows<kSweepEpb>, dim3(grid_of((int64_t)(2 * e.nchunks), kBlock, 65536)), dim3(kBlock),
                       0, s, g.offsets.p, e.rows, e.nchunks, e.chunk_rows.p);
  hipLaunchKernelGGL(k_sv_zero, dim3(grid_of(e.rows, kBlock, 4096)), dim3(kBlock), 0, s, e.seen.p, e.rows);
  for (auto& b : e.fr) hipLaunchKernelGGL(k_sv_zero, dim3(grid_of(e.rows, kBlock, 4096)), dim3(kBlock), 0, s, b.p, e.rows);
  for (auto& b : e.nz) hipLaunchKernelGGL(k_sv_zero, dim3(grid_of(e.words, kBlock, 4096)), dim3(kBlock), 0, s, b.p, e.words);
  for (auto& b : e.nzs) hipLaunchKernelGGL(k_sv_zero, dim3(grid_of(e.words, kBlock, 4096)), dim3(kBlock), 0, s, b.p, e.words);
  G5_HIP(hipGetLastError());
  const int nb = srv->cfg.answer_buffers;
  for (int i = 0; i < nb; ++i) {
    e.bufs.emplace_back(new DevBuf<uint8_t>());
    e.bufs.back()->alloc(e.rows ? e.rows : 1);
    e.free_bufs.push_back(nb - 1 - i);
  }
  for (int r = 0; r < 64; ++r) {
    e.lane_req[r] = -1;
    e.lane_l0[r] = 0;
    e.lane_pulls[r] = 0;
    e.teps_req[r] = -1;
  }
  G5_HIP(hipStreamSynchronize(s));
}

// Consumes the traversed-edge sums of the last retirement: from the level
// whose completion is being processed (its publish carries them; stream order
// puts the retirement's k_sv_teps before that level), or, when no level
// followed the retirement, from a copy once the stream has drained.
static void sv_take_teps(g5_server* srv, Engine& e, bool from_level) {
  if (!e.teps_pending) return;
  if (!from_level) {
    G5_HIP(hipMemcpyAsync(e.h, e.tacc.p, 64 * 8, hipMemcpyDeviceToHost, e.s));
    G5_HIP(hipStreamSynchronize(e.s));
  }
  uint64_t m = e.teps_mask;
  while (m) {
    const int r = __builtin_ctzll(m);
    m &= m - 1;
    Request& q = srv->req[e.teps_req[r]];
    q.edges = (int64_t)((from_level ? e.pub.value(kSvTeps + r) : e.h[r]) / 2);
  }
  e.teps_pending = false;
  e.teps_mask = 0;
}

// Processes a finished level; returns true if anything changed.
static void sv_complete_level(g5_server* srv, Engine& e, double t) {
  sv_take_teps(srv, e, true);
  const uint64_t live = e.pub.value(M_LIVE);
  e.last_fdeg = e.pub.value(M_FDEG);
  e.nzc[e.level & 1] = (int64_t)e.pub.value(M_NZ);  // the level's frontier words now in fr[level & 1]
  e.cooling = 0;  // cleared from seen by the level just finished
  const uint64_t finished = e.active & ~live;
  // each search's own direction for its next level
  const double na = (double)e.ctx->graph.n_active;
  uint64_t m = e.active & live;
  while (m) {
    const int r = __builtin_ctzll(m);
    m &= m - 1;
    const int64_t nq = (int64_t)e.pub.value(M_NUM + r);
    const double fdeg = (double)e.pub.value(M_NUM + 64 + r);
    e.lane_visited[r] += nq;
    e.lane_vdeg[r] += fdeg;
    const int64_t unvisited = (int64_t)na - e.lane_visited[r];
    if (srv->cfg.lane_rule == 3) {
      // pull while the frontier's edges exceed 1/alpha of the edges not yet
      // reached; a search that pulls keeps pulling (its counts are no longer
      // gathered)
      if (!((e.pulled >> r) & 1)) e.lane_fwd[r] = fdeg * srv->cfg.alpha <= (double)e.nnz - e.lane_vdeg[r];
      if (!e.lane_fwd[r]) e.pulled |= UINT64_C(1) << r;
    } else if (srv->cfg.lane_rule == 2) {
      // one direction for all searches, from the frontier union's edges
      e.lane_fwd[r] = (double)e.last_fdeg < (double)e.nnz / srv->cfg.alpha;
    } else if (srv->cfg.lane_rule == 1) {
      // edge rule: pull the searches whose frontier has many edges
      e.lane_fwd[r] = fdeg < (double)e.nnz / srv->cfg.alpha;
    } else if (e.lane_growing[r] && nq > e.lane_prev_nq[r]) {
      if (e.lane_fwd[r] && (double)nq > na / 2000.0) e.lane_fwd[r] = false;
    } else if (!e.lane_fwd[r] && unvisited < (int64_t)(na / 8.0)) {
      e.lane_fwd[r] = true;
      e.lane_growing[r] = false;
    }
    e.lane_prev_nq[r] = nq;
  }
  if (finished) {
    // traversed edges of the finished lanes, before their bits are cleared
    G5_HIP(hipMemsetAsync(e.tacc.p, 0, 64 * 8, e.s));
    hipLaunchKernelGGL(k_sv_teps, dim3(grid_of(e.rows, kBlock, 2048)), dim3(kBlock), 0, e.s, e.seen.p,
                       e.ctx->graph.weight.p, e.rows, finished, e.tacc.p);
    G5_HIP(hipGetLastError());
    e.teps_pending = true;
    e.teps_mask = finished;
    uint64_t m = finished;
    while (m) {
      const int r = __builtin_ctzll(m);
      m &= m - 1;
      const int64_t id = e.lane_req[r];
      Request& q = srv->req[id];
      q.status = SV_DONE;
      q.done = t;
      ++srv->done_count;
      if (id >= srv->run_base) ++srv->run_done;
      ++e.st.completed;
      q.levels = e.level - e.lane_l0[r];
      e.teps_req[r] = id;
      e.lane_req[r] = -1;
      if (!q.keep) {
        e.free_bufs.push_back(q.buf);
        q.buf = -1;
      }
    }
    e.active &= ~finished;
    e.cooling = finished;
  }
}

// Admits waiting queries into free lanes; returns the lanes admitted.  An
// idle graph admits at once.  A graph with searches in flight admits at this
// level boundary, into the lanes and answer arrays that are free,
//   * while every search in flight still pushes: their frontiers are small,
//     and the newcomers reach their own expensive (pulling) levels together
//     with them;
//   * once min(min_admit, free answer arrays) queries can start together (a
//     cohort shares its expensive levels);
//   * once the oldest waiting query has waited max_wait_s.
// Admitting into a graph whose searches already pull gives it searches at
// different depths, and every later level then pulls for some of them.
static uint64_t sv_admit(g5_server* srv, Engine& e, double t) {
  const uint64_t busy_lanes = e.active | e.cooling;
  if (e.waiting.empty() || busy_lanes == ~UINT64_C(0) || e.free_bufs.empty()) return 0;
  const bool in_flight = e.active != 0;
  if (in_flight) {
    bool young = true;
    for (uint64_t x = e.active; x && young; x &= x - 1) young = e.lane_fwd[__builtin_ctzll(x)];
    const int64_t free_lanes = 64 - __builtin_popcountll(busy_lanes);
    const int64_t free_bufs = (int64_t)e.free_bufs.size();
    const int64_t can = std::min<int64_t>({(int64_t)e.waiting.size(), free_lanes, free_bufs});
    const int64_t want = std::min<int64_t>((int64_t)srv->cfg.min_admit, free_bufs);
    const double waited = t - srv->req[e.waiting.front()].arrival;
    if (!young && can < want && waited < srv->cfg.max_wait_s) return 0;
  }
  ServeRoots rs;
  for (int r = 0; r < 64; ++r) {
    rs.root[r] = -1;
    rs.ans[r] = nullptr;
  }
  uint64_t admitted = 0;
  for (int r = 0; r < 64 && !e.waiting.empty() && !e.free_bufs.empty(); ++r) {
    if ((busy_lanes >> r) & 1) continue;
    const int64_t id = e.waiting.front();
    e.waiting.pop_front();
    Request& q = srv->req[id];
    q.lane = r;
    q.buf = e.free_bufs.back();
    e.free_bufs.pop_back();
    q.admit = t;
    q.status = SV_RUNNING;
    e.lane_req[r] = id;
    e.lane_l0[r] = e.level;
    e.lane_fwd[r] = true;
    e.pulled &= ~(UINT64_C(1) << r);
    e.lane_growing[r] = true;
    e.lane_pulls[r] = 0;
    e.lane_prev_nq[r] = 0;
    e.lane_visited[r] = 1;
    e.lane_vdeg[r] = 0;
    uint8_t* ans = e.bufs[q.buf]->p;
    hipLaunchKernelGGL(k_sv_fill_bytes, dim3(grid_of(e.rows, kBlock, 2048)), dim3(kBlock), 0, e.s, ans, e.rows,
                       (uint8_t)0xFF);
    rs.root[r] = q.root;
    rs.ans[r] = ans;
    admitted |= UINT64_C(1) << r;
  }
  if (admitted) {
    hipLaunchKernelGGL(k_sv_admit, dim3(1), dim3(64), 0, e.s, e.ctx->graph.old2new.p, e.rows, rs, e.seen.p,
                       e.fr[e.level & 1].p, e.nz[e.level & 1].p, e.nzs[e.level & 1].p);
    G5_HIP(hipGetLastError());
    e.nzs_lanes[e.level & 1] |= admitted;
    e.nzc[e.level & 1] += __builtin_popcountll(admitted);  // their roots' words join the marks
    e.active |= admitted;
    const int64_t k = __builtin_popcountll(admitted);
    (in_flight ? e.st.admitted_busy : e.st.admitted_idle) += k;
    ++e.admit_events;
  }
  return admitted;
}

// Launches the next level for the active lanes.
static void sv_launch_level(g5_server* srv, Engine& e) {
  Graph& g = e.ctx->graph;
  hipStream_t s = e.s;
  const int cur = e.level & 1;
  const int l = e.level + 1;
  G5_CHECK(l < (1 << 30), "serving: level counter overflow");
  uint64_t* front = e.fr[cur].p;
  uint64_t* next = e.fr[cur ^ 1].p;
  uint64_t* nz_front = e.nz[cur].p;
  uint64_t* nz_next = e.nz[cur ^ 1].p;
  ServeEpi epi;
  uint64_t m = e.active;
  for (int r = 0; r < 64; ++r) {
    epi.ans[r] = nullptr;
    epi.lvl[r] = 0;
  }
  while (m) {
    const int r = __builtin_ctzll(m);
    m &= m - 1;
    const int rel = l - e.lane_l0[r];
    G5_CHECK(rel <= 254, "serving: a query needs more than 254 levels (uint8 level arrays)");
    epi.ans[r] = e.bufs[srv->req[e.lane_req[r]].buf]->p;
    epi.lvl[r] = (uint8_t)rel;
  }
  const uint64_t keep = ~e.cooling;
  // top-down lanes push, bottom-up lanes pull, in the same level
  uint64_t fwd = 0;
  for (int r = 0; r < 64; ++r)
    if (e.lane_fwd[r]) fwd |= UINT64_C(1) << r;
  const uint64_t td = e.active & fwd, bu = e.active & ~fwd;
  unsigned long long* c = e.cnt.p;
  epi.lane_nq = c + M_NUM + 1;
  epi.offsets = g.offsets.p;
  epi.nz_sel = e.nzs[cur ^ 1].p;
  epi.sel = td;
  epi.stat_lanes = srv->cfg.lane_rule == 3 ? ~e.pulled : ~UINT64_C(0);
  // chunks to push: those holding a top-down lane's frontier vertex; the
  // bitmap from the last level's epilogue (and admissions) serves when it
  // was made for exactly these lanes, else it is made from the frontier
  const uint64_t* nz_push = e.nzs[cur].p;
  const bool reselect = td && e.nzs_lanes[cur] != td;
  e.nzs_lanes[cur ^ 1] = td;
  G5_HIP(hipEventRecord(e.ev_start, s));
  if (reselect)
    hipLaunchKernelGGL(k_sv_select, dim3(grid_of(e.words, kBlock / 64, 8192)), dim3(kBlock), 0, s, front, e.rows, td,
                       e.nzs[cur].p);
  G5_HIP(hipMemsetAsync(c, 0, kSvCounters * 8, s));
  ++e.st.levels;
  if (td && bu) ++e.st.mixed_levels; else if (td) ++e.st.push_levels; else ++e.st.pull_levels;
  e.st.lane_levels += __builtin_popcountll(e.active);
  // A push level with few frontier edges and no lane cooling lists the
  // vertices it touches and folds only those (the full fold also clears the
  // cooling lanes from seen).
  const bool listed = td && !bu && !e.cooling && (double)e.last_fdeg * 8.0 < (double)e.rows;
  epi.list = listed ? 1 : 0;
  if (td) {
    // next held the frontier of two levels back, non-zero exactly at the
    // bits of nz_next's last contents: few of them are cleared one by one
    if (e.nzc[cur ^ 1] * 16 < e.rows)
      hipLaunchKernelGGL(k_sv_clear_marked, dim3(grid_of(e.words, kBlock, 4096)), dim3(kBlock), 0, s, nz_next,
                         e.words, e.rows, next);
    else
      hipLaunchKernelGGL(k_sv_zero, dim3(grid_of(e.rows, kBlock, 4096)), dim3(kBlock), 0, s, next, e.rows);
    hipLaunchKernelGGL(k_sv_collect, dim3(grid_of((int64_t)e.nchunks, kBlock, 4096)), dim3(kBlock), 0, s,
                       e.chunk_rows.p, e.nchunks, nz_push, e.chunk_list.p, (unsigned*)(c + M_NUM));
    if (listed) {
      G5_HIP(hipMemsetAsync(nz_next, 0, e.words * 8, s));  // after the clear above read it
      G5_HIP(hipMemsetAsync(epi.nz_sel, 0, e.words * 8, s));
      hipLaunchKernelGGL(k_sv_push_touch, dim3(grid_of((int64_t)e.nchunks, 1, 4096)), dim3(kBlock), 0, s,
                         e.chunk_list.p, (const unsigned*)(c + M_NUM), e.chunk_rows.p, (uint64_t)e.nnz, g.offsets.p,
                         g.cols.p, front, nz_front, next, td, e.touch.p, c + kSvTouch);
    } else {
      hipLaunchKernelGGL(k_sv_push, dim3(grid_of((int64_t)e.nchunks, 1, 4096)), dim3(kBlock), 0, s, e.chunk_list.p,
                         (const unsigned*)(c + M_NUM), e.chunk_rows.p, (uint64_t)e.nnz, g.offsets.p, g.cols.p, front,
                         nz_front, next, td);
    }
  }
  if (listed) {
    hipLaunchKernelGGL(k_sv_finish_list, dim3(grid_of(e.rows / 8 + 1, kBlock, 2048)), dim3(kBlock), 0, s, e.seen.p,
                       next, nz_next, g.offsets.p, e.touch.p, (const unsigned long long*)(c + kSvTouch), c, epi);
  } else if (bu) {
    // a pull of searches that have each pulled twice before finds little to
    // probe (their frontiers have passed the hubs): its pass over every
    // vertex runs on a smaller grid, as the batch's light levels do
    bool light = true;
    for (uint64_t x = bu; x; x &= x - 1) {
      const int r = __builtin_ctzll(x);
      light = light && e.lane_pulls[r] >= 2;
      ++e.lane_pulls[r];
    }
    hipLaunchKernelGGL(k_sv_pull, dim3(grid_of(e.words, kBlock / 64, light ? 2048 : 8192)), dim3(kBlock), 0, s,
                       e.seen.p, front,
                       nz_front, next, nz_next, g.first_nbr.p, g.offsets.p, g.cols.p, e.rows, bu,
                       srv->cfg.pull_filter, c, keep, td ? 1 : 0, epi);
  } else {
    hipLaunchKernelGGL(k_sv_finish, dim3(grid_of(e.words, kBlock / 64, 8192)), dim3(kBlock), 0, s, e.seen.p, next,
                       nz_next, g.offsets.p, e.rows, c, keep, epi);
  }
  G5_HIP(hipGetLastError());
  G5_HIP(hipEventRecord(e.ev_end, s));
  hipLaunchKernelGGL(k_sv_publish, dim3(1), dim3(kBlock), 0, s, c, e.tacc.p, e.pub.d, ++e.pub_seq);
  G5_HIP(hipGetLastError());
  e.level = l;
  e.busy = true;
  ++e.levels_run;
}

// The oldest waiting query of an idle graph, answered on a single-root slot:
// g5_run_bfs's kernels on the slot's state and stream.  The single-root level
// array (n_rows bytes in row order, the root 0, unreached 0xFF) is the answer
// array's format, and its traversed-edge count the engine's.
static void sv_start_solo(g5_server* srv, Engine& e, SoloSlot& sl, double t) {
  const int64_t id = e.waiting.front();
  e.waiting.pop_front();
  Request& q = srv->req[id];
  q.admit = t;
  q.status = SV_RUNNING;
  uint8_t* ans = nullptr;
  if (q.keep) {
    q.buf = e.free_bufs.back();
    e.free_bufs.pop_back();
    ans = e.bufs[q.buf]->p;
  }
  if (sl.th.joinable()) sl.th.join();
  sl.req = id;
  sl.err.clear();
  sl.state.store(1, std::memory_order_release);
  ++e.admit_events;
  ++e.st.admitted_idle;
  SoloSlot* sp = &sl;
  g5_ctx* ctx = e.ctx;
  const int64_t rows = e.rows;
  const int dev = srv->device;
  const double t0 = srv->t0;
  const int64_t root = q.root;
  sl.th = std::thread([sp, ctx, rows, dev, t0, root, ans] {
    try {
      G5_HIP(hipSetDevice(dev));
      bfs_run(ctx, *sp->b, sp->s, root);
      if (ans) {
        G5_HIP(hipMemcpyAsync(ans, sp->b->level.p, (size_t)rows, hipMemcpyDeviceToDevice, sp->s));
        G5_HIP(hipStreamSynchronize(sp->s));
      }
      sp->edges = bfs_teps_edges(ctx, *sp->b, sp->acc, sp->s);
      sp->levels = sp->b->stats.num_levels - 1;
      sp->done = now_s() - t0;
    } catch (const Error& x) {
      sp->err = x.msg.empty() ? "serving: single-root run failed" : x.msg;
    } catch (...) {
      sp->err = "serving: single-root run failed";
    }
    sp->state.store(2, std::memory_order_release);
  });
}

static void sv_finish_solo(g5_server* srv, Engine& e, SoloSlot& sl) {
  sl.th.join();
  sl.state.store(0, std::memory_order_relaxed);
  G5_CHECK(sl.err.empty(), sl.err);
  Request& q = srv->req[sl.req];
  q.status = SV_DONE;
  q.done = sl.done;
  q.edges = sl.edges;
  q.levels = sl.levels;
  ++srv->done_count;
  if (sl.req >= srv->run_base) ++srv->run_done;
  ++e.st.completed;
  ++e.st.solo;
  sl.req = -1;
}

// An idle graph (no search in lanes) with at most cfg.solo waiting queries
// answers the oldest on a free single-root slot, as long as the queue stays
// that short; a longer queue is batched into lanes once the slots are done.
static bool sv_start_solos(g5_server* srv, Engine& e, double t) {
  if (!srv->cfg.solo || e.active) return false;
  bool started = false;
  for (auto& sl : e.solo) {
    if (e.waiting.empty() || e.waiting.size() > (size_t)srv->cfg.solo) break;
    if (srv->req[e.waiting.front()].keep && e.free_bufs.empty()) break;
    if (sl->state.load(std::memory_order_acquire) != 0) continue;
    if (e.solo_running()) ++e.st.solo_overlap;
    sv_start_solo(srv, e, *sl, t);
    started = true;
  }
  return started;
}

// One non-blocking step of an engine; returns true if it made progress.
static bool sv_step(g5_server* srv, Engine& e, double t) {
  bool progress = false, running = false;
  for (auto& sl : e.solo) {
    const int st = sl->state.load(std::memory_order_acquire);
    if (st == 2) {
      sv_finish_solo(srv, e, *sl);
      progress = true;
    } else if (st == 1) {
      running = true;
    }
  }
  // while single-root runs go on, more may start but no query takes a lane
  if (running) return sv_start_solos(srv, e, t) || progress;
  if (e.busy) {
    if (!e.pub.ready(0, kSvPub, e.pub_seq)) {
      // not published yet; a stream that has finished without publishing
      // failed
      const hipError_t q = hipStreamQuery(e.s);
      if (q == hipErrorNotReady) return false;
      G5_HIP(q);
      G5_CHECK(e.pub.ready(0, kSvPub, e.pub_seq), "serving: a level's counters were not published");
    }
    e.busy = false;
    float lms = 0;
    G5_HIP(hipEventSynchronize(e.ev_end));  // recorded before the publish: already complete
    if (hipEventElapsedTime(&lms, e.ev_start, e.ev_end) == hipSuccess) e.st.busy_ms += lms;
    sv_complete_level(srv, e, t);
  }
  if (!e.active && e.cooling && !e.waiting.empty()) {
    // idle engine: clear the cooling lanes directly before reusing them
    hipLaunchKernelGGL(k_sv_clear, dim3(grid_of(e.rows, kBlock, 4096)), dim3(kBlock), 0, e.s, e.seen.p, e.rows,
                       ~e.cooling);
    G5_HIP(hipGetLastError());
    e.cooling = 0;
  }
  if (sv_start_solos(srv, e, t)) return true;
  sv_admit(srv, e, t);
  if (e.active) {
    sv_launch_level(srv, e);
    return true;
  }
  return progress;
}

static void sv_run(g5_server* srv, const g5_query* qs, int64_t nq, double timeout_s, g5_query_result* res,
                   g5_server_report* rep) {
  const int ng = (int)srv->eng.size();
  for (int64_t i = 0; i < nq; ++i) {
    G5_CHECK(qs[i].graph >= 0 && qs[i].graph < ng, "g5_server_run: query graph index out of range");
    G5_CHECK(qs[i].root >= 0 && qs[i].root < srv->eng[qs[i].graph]->n, "g5_server_run: query root out of range");
    G5_CHECK(i == 0 || qs[i].arrival_s >= qs[i - 1].arrival_s, "g5_server_run: arrivals must be non-decreasing");
  }
  // a single-root run left by an earlier run that ended in an error finishes
  // first, so that it does not count as one of this run's answers
  for (auto& e : srv->eng)
    for (auto& sl : e->solo)
      if (sl->state.load(std::memory_order_acquire) != 0) sv_finish_solo(srv, *e, *sl);
  const int64_t base = (int64_t)srv->req.size();
  srv->req.resize(base + nq);
  for (int64_t i = 0; i < nq; ++i) {
    Request& q = srv->req[base + i];
    q.graph = qs[i].graph;
    q.root = qs[i].root;
    q.arrival = qs[i].arrival_s;
    q.keep = qs[i].keep != 0;
  }
  int64_t next_arrival = 0, done = 0;
  // queries left waiting or in lanes by an earlier run that ended in an
  // error are still answered, but only this run's count toward its end
  srv->run_base = base;
  srv->run_done = 0;
  const double t0 = now_s();
  srv->t0 = t0;
  int64_t levels0 = 0, busy0 = 0, events0 = 0;
  for (auto& e : srv->eng) {
    levels0 += e->levels_run;
    busy0 += e->st.admitted_busy;
    events0 += e->admit_events;
  }
  int max_active = 0;
  while (done < nq) {
    const double t = now_s() - t0;
    G5_CHECK_CODE(t <= timeout_s, G5_ERR_LIMIT, "g5_server_run: timed out before every query completed");
    while (next_arrival < nq && srv->req[base + next_arrival].arrival <= t) {
      Request& q = srv->req[base + next_arrival];
      srv->eng[q.graph]->waiting.push_back(base + next_arrival);
      ++next_arrival;
    }
    bool progress = false;
    for (auto& e : srv->eng) {
      progress |= sv_step(srv, *e, t);
      const int a = __builtin_popcountll(e->active);
      if (a > max_active) max_active = a;
    }
    done = srv->run_done;
    if (!progress) {
      bool any_busy = false;
      for (auto& e : srv->eng) any_busy |= e->busy || e->solo_running();
      if (!any_busy && next_arrival < nq) {
        const double wait = srv->req[base + next_arrival].arrival - (now_s() - t0);
        if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 100e-6));
      }
    }
  }
  // drain in-flight work (levels launched for lanes that were all retired) and pending edge counts
  for (auto& e : srv->eng) {
    G5_HIP(hipStreamSynchronize(e->s));
    if (e->busy) {
      e->busy = false;
      sv_complete_level(srv, *e, now_s() - t0);
      G5_HIP(hipStreamSynchronize(e->s));
    }
    sv_take_teps(srv, *e, false);
  }
  const double wall = now_s() - t0;
  int64_t edges = 0, levels1 = 0, busy1 = 0, events1 = 0;
  for (auto& e : srv->eng) {
    levels1 += e->levels_run;
    busy1 += e->st.admitted_busy;
    events1 += e->admit_events;
  }
  for (int64_t i = 0; i < nq; ++i) {
    const Request& q = srv->req[base + i];
    g5_query_result& r = res[i];
    r.request = base + i;
    r.admit_s = q.admit;
    r.done_s = q.done;
    r.edges = q.edges;
    r.levels = q.levels;
    r.status = q.status;
    edges += q.edges > 0 ? q.edges : 0;
  }
  if (rep) {
    rep->wall_s = wall;
    rep->completed = nq;
    rep->edges = edges;
    rep->levels_launched = levels1 - levels0;
    rep->max_active_lanes = max_active;
    rep->reserved = 0;
    rep->admitted_busy = busy1 - busy0;
    rep->admit_events = events1 - events0;
  }
}

}  // namespace g5

using namespace g5;

namespace {
template <class F>
int sv_guarded(g5_server* srv, F&& f) {
  if (!srv) return G5_ERR_ARG;
  try {
    srv->err.clear();
    G5_HIP(hipSetDevice(srv->device));
    f();
    return G5_OK;
  } catch (const Error& e) {
    srv->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    srv->err = "host allocation failed";
    return G5_ERR_MEM;
  } catch (...) {
    srv->err = "unexpected exception";
    return G5_ERR_STATE;
  }
}
}  // namespace

extern "C" {

int g5_server_create(g5_ctx* const* graphs, int num_graphs, const g5_server_config* cfg, g5_server** out) {
  if (!out || !graphs || num_graphs < 1 || num_graphs > 64 || !cfg) return G5_ERR_ARG;
  *out = nullptr;
  if (cfg->answer_buffers < 1 || cfg->answer_buffers > 4096 || !(cfg->alpha > 0) ||
      (cfg->pull_filter != 0 && cfg->pull_filter != 1) || cfg->min_admit < 1 || cfg->min_admit > 64 ||
      cfg->lane_rule < 0 || cfg->lane_rule > 3 ||
      !(cfg->max_wait_s >= 0) || cfg->solo < 0 || cfg->solo > 64 || cfg->solo_slots < 0 || cfg->solo_slots > 8)
    return G5_ERR_ARG;
  for (int i = 0; i < num_graphs; ++i) {
    if (!graphs[i] || graphs[i]->device != graphs[0]->device) return G5_ERR_ARG;
    for (int j = 0; j < i; ++j)
      if (graphs[j] == graphs[i]) return G5_ERR_ARG;
  }
  g5_server* srv = new (std::nothrow) g5_server();
  if (!srv) return G5_ERR_MEM;
  srv->cfg = *cfg;
  srv->device = graphs[0]->device;
  const int rc = sv_guarded(srv, [&] {
    for (int i = 0; i < num_graphs; ++i) {
      srv->eng.emplace_back(new Engine());
      sv_setup_engine(srv, *srv->eng.back(), graphs[i]);
      if (cfg->solo) sv_setup_solo(*srv->eng.back(), std::max(cfg->solo_slots, 1));
    }
  });
  if (rc != G5_OK) {
    // keep the message reachable through a failed create: print nothing, return the code
    delete srv;
    return rc;
  }
  // the contexts refuse g5_construct and g5_destroy while this server exists
  for (auto& e : srv->eng) ++e->ctx->servers;
  *out = srv;
  return G5_OK;
}

void g5_server_destroy(g5_server* srv) {
  if (!srv) return;
  (void)hipSetDevice(srv->device);
  for (auto& e : srv->eng) {
    for (auto& sl : e->solo)
      if (sl->th.joinable()) sl->th.join();  // a single-root run still going
    (void)hipStreamSynchronize(e->s);
    --e->ctx->servers;
  }
  delete srv;
}

const char* g5_server_last_error(const g5_server* srv) { return srv ? srv->err.c_str() : "null server"; }

int g5_server_run(g5_server* srv, const g5_query* queries, int64_t num_queries, double timeout_s,
                  g5_query_result* results, g5_server_report* report) {
  return sv_guarded(srv, [&] {
    G5_CHECK(num_queries >= 0 && (num_queries == 0 || (queries && results)), "g5_server_run: bad arguments");
    G5_CHECK(timeout_s > 0, "g5_server_run: timeout must be positive");
    sv_run(srv, queries, num_queries, timeout_s, results, report);
  });
}

int g5_server_get_levels(g5_server* srv, int64_t request, uint8_t* levels) {
  return sv_guarded(srv, [&] {
    G5_CHECK(levels, "g5_server_get_levels: null output");
    G5_CHECK(request >= 0 && request < (int64_t)srv->req.size(), "g5_server_get_levels: unknown request");
    const Request& q = srv->req[request];
    G5_CHECK(q.status == SV_DONE && q.buf >= 0,
             "g5_server_get_levels: request not complete, not kept (keep = 0) or released");
    Engine& e = *srv->eng[q.graph];
    srv->lv_out.alloc(e.n);
    hipLaunchKernelGGL(k_sv_levels, dim3(grid_of(e.n, kBlock, 65536)), dim3(kBlock), 0, e.s,
                       e.ctx->graph.old2new.p, e.n, e.rows, e.bufs[q.buf]->p, q.root, srv->lv_out.p);
    G5_HIP(hipGetLastError());
    G5_HIP(hipMemcpyAsync(levels, srv->lv_out.p, e.n, hipMemcpyDeviceToHost, e.s));
    G5_HIP(hipStreamSynchronize(e.s));
  });
}

int g5_server_graph_stats_get(g5_server* srv, int graph, g5_server_graph_stats* out) {
  return sv_guarded(srv, [&] {
    G5_CHECK(out && graph >= 0 && graph < (int)srv->eng.size(), "g5_server_graph_stats_get: bad arguments");
    *out = srv->eng[graph]->st;
  });
}

int g5_server_release(g5_server* srv, int64_t request) {
  return sv_guarded(srv, [&] {
    G5_CHECK(request >= 0 && request < (int64_t)srv->req.size(), "g5_server_release: unknown request");
    Request& q = srv->req[request];
    G5_CHECK(q.status == SV_DONE, "g5_server_release: request not complete");
    if (q.buf >= 0) srv->eng[q.graph]->free_bufs.push_back(q.buf);
    q.buf = -1;
    q.status = SV_RELEASED;
  });
}

}  // extern "C"
