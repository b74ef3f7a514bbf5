// g5_msbfs.hip -- batched multi-source BFS: up to 64 roots share one sweep.
//
// Lane r of every 64-bit word belongs to root r.  Per vertex (degree-ordered
// new ids, rows [0, n_rows)):
//   seen[v]        lanes that have reached v
//   fr[l][v]       lanes for which v is at level l (the frontier of level l+1);
//                  kept for every level, so the per-root level arrays are
//                  rebuilt after the run without any extra store in the sweep
// Each level is a push or a pull step, chosen from the edge count of the
// frontier union (the direction-optimizing rule applied to the union):
//   push  the union's rows form a chunked virtual edge list (as in the
//         single-root top-down); target w receives front[u] with one
//         64-bit atomicOr, filtered by a plain load of next[w]; k_ms_finish
//         then masks the lanes new to w and folds them into seen
//   pull  one lane per vertex: the lanes it still lacks (and whose frontier
//         is not empty) are collected by OR-ing the neighbours' frontier
//         words in ascending (highest degree first) order until every needed
//         lane is found or the row ends; long rows are scanned by the whole
//         wave with a wave-wide OR.  A 1-bit-per-vertex "frontier word is not
//         zero" bitmap (n/8 bytes, cache resident) is tested first, so the
//         8-byte frontier word is only read for frontier vertices
// Per-lane levels equal the single-root BFS levels of the same root.
#include "g5_ctx.hpp"
#include "g5_frontier.hpp"
#include "g5_ms_device.hpp"

#include <algorithm>
#include <cstring>

namespace g5 {

constexpr int kMsCounters = M_NUM + 1 + 128;  // level counters, a list count, per-lane counts and degree sums

__global__ void __launch_bounds__(kBlock) k_ms_clear(uint64_t* __restrict__ a, uint64_t* __restrict__ b,
                                                     int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    a[i] = 0;
    b[i] = 0;
  }
}

// Marks every active root in seen and fr0 and counts the level-0 union: one
// wave, lane r takes root r; of the roots sharing a vertex, the lane whose
// atomic OR finds its word still zero counts the vertex.
__global__ void k_ms_roots(const uint32_t* __restrict__ old2new, const int64_t* __restrict__ roots, int nroots,
                           int64_t rows, const uint64_t* __restrict__ offsets, uint64_t* __restrict__ seen,
                           uint64_t* __restrict__ fr0, unsigned long long* __restrict__ c) {
  if (blockIdx.x != 0 || threadIdx.x >= 64) return;
  const int r = threadIdx.x;
  bool act = false, first = false;
  uint64_t deg = 0;
  if (r < nroots && roots[r] >= 0) {  // unused lanes hold -1
    const uint32_t v = old2new[roots[r]];
    act = (int64_t)v < rows;  // a root without edges discovers nothing
    if (act) {
      const unsigned long long bit = 1ull << r;
      atomicOr((unsigned long long*)(seen + v), bit);
      first = atomicOr((unsigned long long*)(fr0 + v), bit) == 0;
      if (first) deg = offsets[v + 1] - offsets[v];
    }
  }
  const uint64_t live = __ballot(act);
  const uint64_t nz = (uint64_t)__popcll(__ballot(first));
  const uint64_t fdeg = wave_sum(deg);
  if (r == 0) {
    c[M_LIVE] = live;
    c[M_PAIRS] = 0;
    c[M_NZ] = nz;
    c[M_FDEG] = fdeg;
  }
}

// Level-0 nonzero bitmap: the root vertices (lane r, root r).
__global__ void k_ms_roots_nz(const uint32_t* __restrict__ old2new, const int64_t* __restrict__ roots, int nroots,
                              int64_t rows, uint64_t* __restrict__ nz0) {
  const int r = threadIdx.x;
  if (blockIdx.x != 0 || r >= nroots || roots[r] < 0) return;
  const uint32_t v = old2new[roots[r]];
  if ((int64_t)v < rows) atomicOr((unsigned long long*)(nz0 + (v >> 6)), 1ull << (v & 63));
}

__global__ void __launch_bounds__(kBlock) k_ms_push(const uint32_t* __restrict__ queue, int64_t nf,
                                                    const uint64_t* __restrict__ prefix,
                                                    const uint32_t* __restrict__ chunk_rows,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ cols,
                                                    const uint64_t* __restrict__ front, uint64_t* next,
                                                    uint64_t lanes) {
  // front[u] (its lanes in `lanes`) goes to every target; lanes the target
  // has seen are dropped when k_ms_finish (or the merging pull) folds next
  // into seen.  prefix[nf] is the edge count: nf may be an upper bound whose
  // tail entries have degree 0.
  __shared__ ChunkRows cr;
  const uint64_t E = prefix[nf];
  for (uint64_t chunk = blockIdx.x; chunk * kEpb < E; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kEpb;
    const uint64_t e1 = (e0 + kEpb < E) ? e0 + kEpb : E;
    const int cnt = stage_chunk(chunk, chunk_rows, queue, prefix, offsets, cr);
    walk_strided(e0, e1, cnt, cr, cols, [&](uint32_t w, uint32_t u) {
      const uint64_t m = front[u] & lanes;
      if (m & ~__atomic_load_n(next + w, __ATOMIC_RELAXED))
        atomicOr((unsigned long long*)(next + w), (unsigned long long)m);
    });
    __syncthreads();
  }
}

// k_ms_push that also lists the vertices it touches: the one atomicOr that
// finds a target's word still zero (next was cleared for the level) appends
// the target, staged in LDS and appended with one global atomic per chunk.
__global__ void __launch_bounds__(kBlock) k_ms_push_touch(const uint32_t* __restrict__ queue, int64_t nf,
                                                          const uint64_t* __restrict__ prefix,
                                                          const uint32_t* __restrict__ chunk_rows,
                                                          const uint64_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ cols,
                                                          const uint64_t* __restrict__ front, uint64_t* next,
                                                          uint32_t* __restrict__ list,
                                                          unsigned long long* __restrict__ list_count) {
  __shared__ ChunkRows cr;
  __shared__ uint32_t s_q[kEpb];
  __shared__ uint32_t s_qn;
  __shared__ uint64_t s_qbase;
  const uint64_t E = prefix[nf];
  for (uint64_t chunk = blockIdx.x; chunk * kEpb < E; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kEpb;
    const uint64_t e1 = (e0 + kEpb < E) ? e0 + kEpb : E;
    if (threadIdx.x == 0) s_qn = 0;
    const int cnt = stage_chunk(chunk, chunk_rows, queue, prefix, offsets, cr);
    walk_strided(e0, e1, cnt, cr, cols, [&](uint32_t w, uint32_t u) {
      const uint64_t m = front[u];
      if (m & ~__atomic_load_n(next + w, __ATOMIC_RELAXED)) {
        const uint64_t old = atomicOr((unsigned long long*)(next + w), (unsigned long long)m);
        if (old == 0) s_q[atomicAdd(&s_qn, 1u)] = w;
      }
    });
    __syncthreads();
    if (threadIdx.x == 0) s_qbase = s_qn ? atomicAdd(list_count, (unsigned long long)s_qn) : 0;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < s_qn; j += kBlock) list[s_qbase + j] = s_q[j];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) k_ms_finish_list(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                                           uint64_t* __restrict__ nz_next,
                                                           const uint64_t* __restrict__ offsets,
                                                           const uint32_t* __restrict__ list,
                                                           const unsigned long long* __restrict__ list_count,
                                                           unsigned long long* __restrict__ c, LaneEpi epi_in,
                                                           int lane_stats) {
  const uint64_t n = *list_count;
  if (lane_stats) {
    const LaneEpi epi = epi_in;  // a copy in registers: the epilogue accumulates per thread
    ms_finish_list_body(seen, next, nz_next, offsets, list, n, c, blockIdx.x, gridDim.x, epi);
  } else {
    ms_finish_list_body(seen, next, nz_next, offsets, list, n, c, blockIdx.x, gridDim.x);
  }
}

__global__ void __launch_bounds__(kBlock) k_ms_finish(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                                      uint64_t* __restrict__ nz_next,
                                                      const uint64_t* __restrict__ offsets, int64_t rows,
                                                      unsigned long long* __restrict__ c) {
  ms_finish_body(seen, next, nz_next, offsets, rows, c, blockIdx.x, gridDim.x);
}

template <int kU, int kW>
__global__ void __launch_bounds__(kBlock) k_ms_pull(uint64_t* __restrict__ seen, const uint64_t* __restrict__ front,
                                                    const uint64_t* __restrict__ nz_front, uint64_t* __restrict__ next,
                                                    uint64_t* __restrict__ nz_next,
                                                    const uint32_t* __restrict__ first_nbr,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ cols, int64_t rows,
                                                    uint64_t live_in, int mode, const uint64_t* __restrict__ pack,
                                                    const uint64_t* __restrict__ compact, uint32_t filter_from,
                                                    unsigned long long* __restrict__ c, uint64_t* __restrict__ miss) {
  const FrontView fv{front, nz_front, pack, compact, mode, filter_from};
  ms_pull_body<NoEpilogue, kU, kW>(seen, fv, next, nz_next, first_nbr, offsets, cols, rows, live_in, c, blockIdx.x,
                                   gridDim.x, ~UINT64_C(0), NoEpilogue(), false, miss);
}

// The pull of a tail level over the listed vertices (see ms_pull_list_body).
__global__ void __launch_bounds__(kBlock) k_ms_pull_list(uint64_t* __restrict__ seen, const uint64_t* __restrict__ front,
                                                         uint64_t* __restrict__ next, uint64_t* __restrict__ nz_next,
                                                         const uint32_t* __restrict__ first_nbr,
                                                         const uint64_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ cols,
                                                         const uint32_t* __restrict__ list,
                                                         const uint64_t* __restrict__ list_count, uint64_t live_in,
                                                         uint64_t* __restrict__ miss,
                                                         unsigned long long* __restrict__ c) {
  const FrontView fv{front, nullptr, nullptr, nullptr, kFrontPlain, 0};
  ms_pull_list_body(seen, fv, next, nz_next, first_nbr, offsets, cols, list, *list_count, live_in, miss, c,
                    blockIdx.x, gridDim.x);
}

// --- per-root directions (lane_rule 1) ---
__global__ void __launch_bounds__(kBlock) k_ms_finish_lanes(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                                            uint64_t* __restrict__ nz_next,
                                                            const uint64_t* __restrict__ offsets, int64_t rows,
                                                            unsigned long long* __restrict__ c, LaneEpi epi_in) {
  const LaneEpi epi = epi_in;  // a copy in registers: the epilogue accumulates per thread
  ms_finish_body(seen, next, nz_next, offsets, rows, c, blockIdx.x, gridDim.x, ~UINT64_C(0), epi);
}

// Pull for the lanes in live_in; with merge, next already holds this level's
// pushed lanes, which are folded in.
template <int kU, int kW>
__global__ void __launch_bounds__(kBlock) k_ms_pull_lanes(uint64_t* __restrict__ seen, const uint64_t* __restrict__ front,
                                                          const uint64_t* __restrict__ nz_front, uint64_t* __restrict__ next,
                                                          uint64_t* __restrict__ nz_next,
                                                          const uint32_t* __restrict__ first_nbr,
                                                          const uint64_t* __restrict__ offsets,
                                                          const uint32_t* __restrict__ cols, int64_t rows,
                                                          uint64_t live_in, int filter, int merge,
                                                          unsigned long long* __restrict__ c, LaneEpi epi_in) {
  const LaneEpi epi = epi_in;  // a copy in registers: the epilogue accumulates per thread
  const FrontView fv{front, nz_front, nullptr, nullptr, filter ? kFrontFilter : kFrontPlain, 0};
  ms_pull_body<LaneEpi, kU, kW>(seen, fv, next, nz_next, first_nbr, offsets, cols, rows, live_in, c, blockIdx.x,
                                gridDim.x, ~UINT64_C(0), epi, merge != 0);
}

// One bit per vertex whose frontier word has one of `lanes`.
__global__ void __launch_bounds__(kBlock) k_ms_select(const uint64_t* __restrict__ front, int64_t rows, uint64_t lanes,
                                                      uint64_t* __restrict__ nz_sel) {
  ms_select_body(front, rows, lanes, nz_sel, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_ms_collect(const uint32_t* __restrict__ chunk_rows, uint64_t nchunks,
                                                       const uint64_t* __restrict__ nz_sel, uint32_t* __restrict__ list,
                                                       unsigned* count) {
  ms_collect_chunks_body(chunk_rows, nchunks, nz_sel, list, count, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_ms_push_list(const uint32_t* __restrict__ list, const unsigned* count,
                                                         const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                                         const uint64_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ cols,
                                                         const uint64_t* __restrict__ front,
                                                         const uint64_t* __restrict__ nz_sel, uint64_t* next,
                                                         uint64_t lanes) {
  __shared__ SweepLds L;
  ms_push_list_body(list, count, chunk_rows, nnz, offsets, cols, front, nz_sel, next, L, blockIdx.x, gridDim.x, lanes);
}

__global__ void __launch_bounds__(kBlock) k_ms_sweep_clear(const uint32_t* __restrict__ chunk_rows,
                                                           uint64_t nchunks, const uint64_t* __restrict__ offsets,
                                                           uint64_t* __restrict__ next) {
  ms_sweep_clear_body(chunk_rows, nchunks, offsets, next, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_ms_sweep(const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                                     const uint64_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ cols,
                                                     const uint64_t* __restrict__ front,
                                                     const uint64_t* __restrict__ nz_front,
                                                     uint64_t* __restrict__ next) {
  __shared__ SweepLds L;
  ms_sweep_body(chunk_rows, nnz, offsets, cols, front, nz_front, next, L, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Persistent level-fused batch: one launch runs every level, the grid meeting
// at a barrier between the phases of a level.  All blocks read the same level
// counters after a barrier and so take the same direction decision; the loop
// ends when no lane discovers anything, when the preallocated frontier-word
// arrays run out (the host loop continues from there), or when a barrier wait
// exceeds its deadline (then every block leaves and the host reports it).
// ---------------------------------------------------------------------------
// Two-level arrival: workgroups count in on one of up to kBarMaxGroups group
// counters (kBarGroup each, one 128-byte line apiece), the last of a group
// counts in on the top counter, and the last group flips the generation that
// everyone waits on; this keeps the atomics on any one address to a few tens
// per barrier instead of one per workgroup.
// Tail pull levels go list-driven once fewer than rows / kListBelow (root,
// vertex) pairs are unreached.
constexpr double kListBelow = 16.0;

// A push level lists the vertices it touches when its frontier edges times
// this are fewer than the rows (the touched vertices are at most the edges).
constexpr double kTouchBelow = 8.0;

// Pull levels with fewer than 1/kLightBelow of the (root, vertex) pairs
// unreached run on at most kLightGrid workgroups: their pass over every
// vertex finds little to probe, and at SCALE 28 it took 0.48 ms on 2,048
// workgroups against 0.86 ms on 16,384 and 1.59 ms on 32,768, while the
// mixed level gains from the larger grids (26.2 ms on 2,048, 24.6 ms on
// 32,768).
constexpr double kLightBelow = 8.0;
constexpr int64_t kLightGrid = 2048;

constexpr unsigned kBarGroup = 16;
constexpr unsigned kBarMaxGroups = 1024;
struct GridBarrier {
  unsigned count;
  unsigned gen;
  unsigned abort;
  unsigned pad[29];
  unsigned group[kBarMaxGroups][32];
};

// Returns false (in every thread of the block) once any block has given up.
__device__ __forceinline__ bool grid_barrier(GridBarrier* b, unsigned nblk, long long deadline) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned grp = blockIdx.x / kBarGroup;
    const unsigned ngroups = (nblk + kBarGroup - 1) / kBarGroup;
    const unsigned gsize = (grp + 1 < ngroups) ? kBarGroup : nblk - grp * kBarGroup;
    bool last = false;
    if (__hip_atomic_fetch_add(&b->group[grp][0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gsize - 1) {
      __hip_atomic_store(&b->group[grp][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == ngroups - 1;
    }
    if (last) {
      __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // relaxed polls (an acquire per poll would invalidate the caches the
      // still-working workgroups depend on); one acquire fence after
      while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            wall_clock64() > deadline) {
          __hip_atomic_store(&b->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

struct MsPersistArgs {
  const uint64_t* offsets;
  const uint32_t* cols;
  const uint32_t* first_nbr;
  const uint32_t* sweep_rows;  // first/last row per sweep chunk
  int64_t rows;
  uint64_t nnz;
  uint64_t sweep_chunks;
  uint64_t* seen;
  uint64_t* const* fr;          // frontier words of levels 0 .. max_levels-1
  uint64_t* nz0;
  uint64_t* nz1;
  unsigned long long* cnt;      // [max_levels][M_NUM]; level 0 filled by k_ms_roots, rest zero
  long long* ts;                // [max_levels] wall clock at the end of each level
  int* dir;                     // [max_levels] level kind (1 push, 0 pull, 2 sweep)
  int* done_levels;             // last level completed
  GridBarrier* bar;
  uint32_t* chunk_list;         // sweep chunks holding a frontier row (push levels)
  unsigned* chunk_count;
  double push_limit;            // nnz / alpha
  double sweep_below;           // sweep while reached pairs < sweep_below x roots x rows
  int pull_filter;              // pull reads a frontier word only after its nonzero bit
  double pull_unseen_frac;      // pull whenever fewer than this fraction of pairs are unreached
  int max_levels;
  long long timeout_ticks;      // per-barrier wait bound (wall clock ticks)
};

// kWaves: waves per SIMD the compiler must fit (register budget 512 / kWaves).
template <int kWaves>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWaves, kWaves)))
k_ms_persistent(MsPersistArgs a) {
  __shared__ SweepLds L;
  const unsigned bid = blockIdx.x, nblk = gridDim.x;
  const int64_t words = (a.rows + 63) / 64;
  if (bid == 0 && threadIdx.x == 0) a.ts[0] = wall_clock64();
  uint64_t live = __hip_atomic_load(&a.cnt[M_LIVE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double roots_live = (double)__popcll(live);
  const double sweep_limit = a.sweep_below * roots_live * (double)a.rows;
  const double total_pairs = roots_live * (double)a.rows;
  double reached = roots_live;
  int l = 0;
  while (live && l + 1 < a.max_levels) {
    const unsigned long long* cp = a.cnt + (size_t)l * M_NUM;
    const uint64_t fdeg = __hip_atomic_load(&cp[M_FDEG], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ++l;
    unsigned long long* c = a.cnt + (size_t)l * M_NUM;
    const uint64_t* front = a.fr[l - 1];
    uint64_t* next = a.fr[l];
    const uint64_t* nz_front = ((l - 1) & 1) ? a.nz1 : a.nz0;
    uint64_t* nz_next = (l & 1) ? a.nz1 : a.nz0;
    // push a small frontier union, unless nearly every pair is reached (then
    // a pull touches little beyond the seen words)
    const bool push = (double)fdeg < a.push_limit && total_pairs - reached >= a.pull_unseen_frac * total_pairs;
    const bool sweep = !push && a.sweep_chunks && reached < sweep_limit;
    const long long deadline = wall_clock64() + a.timeout_ticks;
    if (push) {
      ms_zero_body(next, a.rows, bid, nblk);
      if (bid == 0 && threadIdx.x == 0) *a.chunk_count = 0;
      if (!grid_barrier(a.bar, nblk, deadline)) return;
      ms_collect_chunks_body(a.sweep_rows, a.sweep_chunks, nz_front, a.chunk_list, a.chunk_count, bid, nblk);
      if (!grid_barrier(a.bar, nblk, deadline)) return;
      ms_push_list_body(a.chunk_list, a.chunk_count, a.sweep_rows, a.nnz, a.offsets, a.cols, front, nz_front,
                        next, L, bid, nblk);
      if (!grid_barrier(a.bar, nblk, deadline)) return;
      ms_finish_body(a.seen, next, nz_next, a.offsets, a.rows, c, bid, nblk);
    } else if (sweep) {
      ms_sweep_clear_body(a.sweep_rows, a.sweep_chunks, a.offsets, next, bid, nblk);
      if (!grid_barrier(a.bar, nblk, deadline)) return;
      ms_sweep_body(a.sweep_rows, a.nnz, a.offsets, a.cols, front, nz_front, next, L, bid, nblk);
      if (!grid_barrier(a.bar, nblk, deadline)) return;
      ms_finish_body(a.seen, next, nz_next, a.offsets, a.rows, c, bid, nblk);
    } else {
      const FrontView fv{front, nz_front, nullptr, nullptr, a.pull_filter == 1 ? kFrontFilter : kFrontPlain, 0};
      ms_pull_body(a.seen, fv, next, nz_next, a.first_nbr, a.offsets, a.cols, a.rows, live, c, bid, nblk);
    }
    if (!grid_barrier(a.bar, nblk, deadline)) return;
    live = __hip_atomic_load(&c[M_LIVE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    reached += (double)__hip_atomic_load(&c[M_PAIRS], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bid == 0 && threadIdx.x == 0) {
      a.ts[l] = wall_clock64();
      a.dir[l] = push ? 1 : (sweep ? 2 : 0);
      *a.done_levels = l;
    }
  }
  (void)words;
}

// Zero fill (graph-capturable, unlike a memset node).
__global__ void __launch_bounds__(kBlock) k_ms_fill(uint64_t* __restrict__ a, int64_t n) {
  ms_zero_body(a, n, blockIdx.x, gridDim.x);
}

// Per-root level of every original vertex: the first level whose frontier
// word holds the lane.
__global__ void k_ms_levels(const uint32_t* __restrict__ old2new, int64_t n, int64_t rows,
                            const uint64_t* const* __restrict__ fr, int nlevels, int lane, int64_t root,
                            uint8_t* __restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t nv = old2new[v];
    uint8_t l = 0xFF;
    if (v == root)
      l = 0;  // also a root without edges, which is no row of the CSR
    else if ((int64_t)nv < rows)
      for (int i = 0; i < nlevels; ++i)
        if ((fr[i][nv] >> lane) & 1) {
          l = (uint8_t)i;
          break;
        }
    out[v] = l;
  }
}

// Traversed-edge numerators of all lanes: sum of weight over the vertices
// each lane reached.  Vertices reached by every lane go to one accumulator.
__global__ void __launch_bounds__(kBlock) k_ms_teps(const uint64_t* __restrict__ seen,
                                                    const uint32_t* __restrict__ weight, int64_t rows,
                                                    uint64_t all, unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long s_acc[65];
  if (threadIdx.x < 65) s_acc[threadIdx.x] = 0;
  __syncthreads();
  uint64_t full = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += (int64_t)gridDim.x * kBlock) {
    const uint64_t s = seen[v];
    const uint64_t w = weight[v];
    if (s == all) {
      full += w;
    } else {
      uint64_t x = s;
      while (x) {
        atomicAdd(&s_acc[__builtin_ctzll(x)], (unsigned long long)w);
        x &= x - 1;
      }
    }
  }
  full = wave_sum(full);
  if ((threadIdx.x & 63) == 0 && full) atomicAdd(&s_acc[64], (unsigned long long)full);
  __syncthreads();
  if (threadIdx.x < 65 && s_acc[threadIdx.x]) atomicAdd(acc + threadIdx.x, s_acc[threadIdx.x]);
}

// --- batched byte model (evaluated after the run from the per-level words) ---
// level l: 24 n + sum over vertices with a needed pulled lane of (8 + 12 k),
// k = probes until every needed lane is found (the row length if not), +
// sum over vertices whose frontier word holds a pushed lane of (8 + 12 deg):
// the word, then per entry the column and the target's next word.  With one
// direction for all roots every level is costed as a pull of all lanes.
__global__ void __launch_bounds__(kBlock) k_ms_model(const uint64_t* const* __restrict__ fr, int l,
                                                     uint64_t lanes, uint64_t push_lanes,
                                                     const uint64_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ cols, int64_t rows,
                                                     unsigned long long* __restrict__ acc) {
  uint64_t b = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += (int64_t)gridDim.x * kBlock) {
    uint64_t seen_before = 0;
    for (int i = 0; i < l; ++i) seen_before |= fr[i][v];
    const uint64_t need = ~seen_before & lanes;
    const uint64_t s = offsets[v], e = offsets[v + 1];
    if (fr[l - 1][v] & push_lanes) b += 8 + 12 * (e - s);
    if (!need || s == e) continue;
    uint64_t a = 0, k = 0;
    for (uint64_t j = s; j < e; ++j) {
      ++k;
      a |= fr[l - 1][cols[j]];
      if ((a & need) == need) break;
    }
    b += 8 + 12 * k;
  }
  block_add(b, acc);
}

// ---------------------------------------------------------------------------
// Host
// ---------------------------------------------------------------------------
// Frontier levels preallocated by g5_batch_prepare (a SCALE 28 batch takes 9).
constexpr int kMsPreLevels = 16;

// The compact frontier copy of the pull (compact_below > 0).
void ms_alloc_compact(g5_ctx* ctx) {
  MsState& m = ctx->ms;
  const int64_t rows = ctx->graph.n_rows;
  m.pack.alloc(2 * ((rows + 63) / 64));
  m.compact.alloc(rows);
}

void ms_prepare(g5_ctx* ctx) {
  Graph& g = ctx->graph;
  G5_CHECK(g.built, "g5_batch_prepare: no graph (call g5_construct first)");
  MsState& m = ctx->ms;
  m.release();
  const int64_t rows = g.n_rows;
  m.seen.alloc(rows);
  for (auto& b : m.nz) b.alloc((rows + 63) / 64);
  m.queue.alloc(rows);
  m.prefix.alloc(rows + 1);
  m.chunk_rows.alloc(2 * ((uint64_t)g.nnz / kEpb + 2));
  m.scan_tmp.alloc(scan_tmp_elems(rows + 1));
  m.counters.alloc(kMsCounters);
  m.roots.alloc(64);
  m.frp.alloc(G5_MAX_LEVELS);
  if (g.reorder_mode != 0 && g.nnz > 0) {
    // row range of every sweep chunk of the CSR (static for the graph)
    m.sweep_chunks = ((uint64_t)g.nnz + kSweepEpb - 1) / kSweepEpb;
    m.sweep_rows.alloc(2 * m.sweep_chunks);
    hipLaunchKernelGGL(k_chunk_rows<kSweepEpb>, dim3(grid_of((int64_t)(2 * m.sweep_chunks), kBlock, 65536)),
                       dim3(kBlock), 0, ctx->stream, g.offsets.p, rows, m.sweep_chunks, m.sweep_rows.p);
    G5_HIP(hipGetLastError());
  }
  // Every buffer a level may need, sized here so that no batch allocates
  // (the in-loop alloc calls below are then no-ops):
  //   chunk_list / nz_sel  the mixed level's chunk list and selection bitmap
  //   touch                a listed push level has fewer than rows / kTouchBelow frontier edges
  //   miss / ulist         a list-driven tail level follows a level with fewer than
  //                        rows / kListBelow unreached (root, vertex) pairs
  //   fr[0, kMsPreLevels)  the frontier words of the first levels (deeper
  //                        searches add levels as they reach them)
  const int64_t words = (rows + 63) / 64;
  if (m.sweep_chunks) m.chunk_list.alloc(m.sweep_chunks);
  m.nz_sel.alloc(words);
  m.touch.alloc((size_t)((double)rows / kTouchBelow) + 2);
  m.miss.alloc(words);
  m.ulist.alloc((size_t)((double)rows / kListBelow) + 2);
  m.ulist_count.alloc(1);
  for (int l = 0; l < kMsPreLevels; ++l) {
    while ((int)m.fr.size() <= l) m.fr.emplace_back(new DevBuf<uint64_t>());
    m.fr[l]->alloc(rows);
    while ((int)m.fr_dirty.size() <= l) m.fr_dirty.emplace_back(new DevBuf<uint64_t>());
    m.fr_dirty[l]->alloc(words);
  }
  m.fr_exact.assign(m.fr.size(), 0);
  m.fr_count.assign(m.fr.size(), 0);
  if (m.compact_below > 0) ms_alloc_compact(ctx);
  G5_HIP(hipHostMalloc((void**)&m.h_counters, kMsCounters * sizeof(uint64_t), hipHostMallocDefault));
  m.pub.alloc(kMsCounters);
  m.pub_seq = 0;
  G5_HIP(hipEventCreate(&m.ev0));
  G5_HIP(hipEventCreate(&m.ev1));
  m.lev_begin.assign(G5_MAX_LEVELS, nullptr);
  m.lev_end.assign(G5_MAX_LEVELS, nullptr);
  for (int i = 0; i < G5_MAX_LEVELS; ++i) {
    G5_HIP(hipEventCreate(&m.lev_begin[i]));
    G5_HIP(hipEventCreate(&m.lev_end[i]));
  }
  m.prepared = true;
}

// Level end: the first n_pub counters the host decides the next level with
// are published to coherent host memory, each tagged with the level's
// sequence number, and every counter is cleared for the next level (so no
// fill starts a level).
__global__ void k_ms_publish(unsigned long long* __restrict__ c, int n_pub, Published* __restrict__ p,
                             uint64_t seq) {
  for (int t = threadIdx.x; t < kMsCounters; t += blockDim.x) {
    const uint64_t v = c[t];
    c[t] = 0;
    if (t < n_pub) publish(p + t, v, seq);
  }
}

static void ms_publish_wait(MsState& m, int n_pub, hipStream_t s) {
  ++m.pub_seq;
  hipLaunchKernelGGL(k_ms_publish, dim3(1), dim3(kBlock), 0, s, (unsigned long long*)m.counters.p, n_pub, m.pub.d,
                     m.pub_seq);
  G5_HIP(hipGetLastError());
  m.pub.wait(0, n_pub, m.pub_seq, s, "g5_run_bfs_batch: level counters were not published");
  for (int i = 0; i < n_pub; ++i) m.h_counters[i] = m.pub.value(i);
}

// Frontier words of level l, allocated on first use and kept for reuse.
static uint64_t* ms_level(MsState& m, int l, int64_t rows) {
  while ((int)m.fr.size() <= l) m.fr.emplace_back(new DevBuf<uint64_t>());
  const uint64_t* before = m.fr[l]->p;
  m.fr[l]->alloc(rows);
  if ((int)m.fr_exact.size() <= l) {
    m.fr_exact.resize(l + 1, 0);
    m.fr_count.resize(l + 1, 0);
  }
  if (m.fr[l]->p != before) m.fr_exact[l] = 0;  // new memory: contents unknown
  return m.fr[l]->p;
}

// Zeroes level l's frontier words before a level that writes them sparsely
// (a push's atomic ORs, a listed fold or pull): only the words its last run
// left non-zero when those are known, else the whole array.
// One bitmap word per lane (the marks are read coalesced); a lane zeroes the
// frontier words of its word's set bits (used for sparse marks only).
__global__ void __launch_bounds__(kBlock) k_ms_clear_marked(const uint64_t* __restrict__ marked, int64_t words,
                                                            int64_t rows, uint64_t* __restrict__ fr) {
  for (int64_t wd = (int64_t)blockIdx.x * kBlock + threadIdx.x; wd < words; wd += (int64_t)gridDim.x * kBlock) {
    for (uint64_t b = marked[wd]; b; b &= b - 1) {
      const int64_t v = wd * 64 + __builtin_ctzll(b);
      if (v < rows) fr[v] = 0;
    }
  }
}

static void ms_clear_frontier(MsState& m, int l, int64_t rows, hipStream_t s, bool marked) {
  const int64_t words = (rows + 63) / 64;
  // a dense mark set is cleared faster by a plain fill
  if (marked && m.fr_count[l] * 16 < rows) {
    hipLaunchKernelGGL(k_ms_clear_marked, dim3(grid_of(words, kBlock, 8192)), dim3(kBlock), 0, s,
                       m.fr_dirty[l]->p, words, rows, m.fr[l]->p);
    G5_HIP(hipGetLastError());
  } else {
    G5_HIP(hipMemsetAsync(m.fr[l]->p, 0, rows * 8, s));
  }
}

// Level l's nonzero bitmap (one bit per vertex whose frontier word is not
// zero), kept per level: it marks the words the level's next run clears.
static uint64_t* ms_marks(MsState& m, int l, int64_t words) {
  while ((int)m.fr_dirty.size() <= l) m.fr_dirty.emplace_back(new DevBuf<uint64_t>());
  m.fr_dirty[l]->alloc(words);
  return m.fr_dirty[l]->p;
}

// Host-driven levels from level l0 on (counters of level l0 - 1 in live /
// nz / fdeg, reached pairs so far in reached): one or more launches per
// level, whose counters the host reads from coherent host memory (the
// counters are clear when it is called).
static void ms_host_levels(g5_ctx* ctx, int l0, uint64_t live, uint64_t nz, uint64_t fdeg, double reached,
                           double roots_live) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  hipStream_t s = ctx->stream;
  const int64_t rows = g.n_rows;
  const int64_t words = (rows + 63) / 64;
  unsigned long long* cnt = (unsigned long long*)m.counters.p;
  g5_batch_stats& st = m.stats;
  const double push_limit = (double)g.nnz / m.alpha;
  const double sweep_limit = m.sweep_below * roots_live * (double)rows;
  const double total_pairs = roots_live * (double)rows;
  const double unseen_frac = m.beta > 0 ? 1.0 / m.beta : 0.0;
  // per-root rule (from the first level, reorder modes 1 and 2): each root's
  // frontier edges and the edges of the vertices it has reached
  const bool lanes_mode = m.lane_rule == 1 && m.sweep_chunks && l0 == 1;
  m.level_masks = lanes_mode;
  double lane_fd[64] = {}, lane_vd[64] = {};
  uint64_t pulled = 0;  // roots that have pulled: they pull from then on
  const uint64_t* h_lane = m.h_counters + M_NUM + 1;
  unsigned* d_list_count = (unsigned*)(cnt + M_NUM);
  bool have_miss = false;  // m.miss holds the vertices still missing a live root
  int l = l0 - 1;
  while (live) {
    ++l;
    bool list_level = false;  // this level writes m.miss
    G5_CHECK(l < G5_MAX_LEVELS - 1, "g5_run_bfs_batch: more than 254 levels (uint8 level arrays)");
    uint64_t* next = ms_level(m, l, rows);
    const uint64_t* front = m.fr[l - 1]->p;  // after ms_level: the vector may have grown
    // whether the marks describe this level's frontier words; unset until
    // the level has been enqueued whole (an error midway leaves them unset)
    const bool marked = m.fr_exact[l] != 0;
    m.fr_exact[l] = 0;
    // each level's nonzero bitmap stays with its frontier words (fr_dirty)
    // for the next batch's clear of them
    uint64_t* nz_next = ms_marks(m, l, words);
    const uint64_t* nz_front = m.fr_dirty[l - 1]->p;
    bool push = (double)fdeg < push_limit && total_pairs - reached >= unseen_frac * total_pairs;
    bool sweep = !push && m.sweep_chunks && reached < sweep_limit;
    uint64_t td = 0, bu = 0;
    bool lane_stats = false;  // this level's epilogue gathers per-root statistics
    if (lanes_mode) {
      // pull when the frontier's edges exceed 1/alpha of the edges not yet
      // reached (the roots' own level pushes; all pull in the beta tail); a
      // root that has pulled keeps pulling, so once every live root pulls
      // the level is a plain pull with no per-root statistics
      const bool tail = total_pairs - reached < unseen_frac * total_pairs;
      for (int r = 0; r < 64; ++r) {
        const uint64_t bit = UINT64_C(1) << r;
        if (!(live & bit)) continue;
        const double unreached = (double)g.nnz - lane_vd[r];
        const bool pull_r = tail || (pulled & bit) || (l > 1 && lane_fd[r] * m.alpha > unreached);
        (pull_r ? bu : td) |= bit;
      }
      pulled |= bu;
      push = bu == 0;
      sweep = false;
      lane_stats = td != 0;
    }
    m.level_push[l] = push ? live : td;
    m.level_pull[l] = push ? 0 : (lanes_mode ? bu : live);
    G5_HIP(hipEventRecord(m.lev_begin[l], s));
    // statistics for the roots still choosing (a root that pulled keeps pulling)
    const LaneEpi lepi{cnt + M_NUM + 1, g.offsets.p, lanes_mode ? td : live};
    if (lanes_mode && td && bu) {
      // the pushing roots first, then the pull of the others, merging the
      // pushed lanes
      m.nz_sel.alloc(words);
      hipLaunchKernelGGL(k_ms_select, dim3(grid_of(words, kBlock / 64, 8192)), dim3(kBlock), 0, s, front, rows, td,
                         m.nz_sel.p);
      ms_clear_frontier(m, l, rows, s, marked);
      if (m.mixed_push == 1) {
        // the pushing roots' frontier vertices as a queue (the count stays on
        // the device), their rows as one edge-balanced virtual edge list; the
        // host sizes the scans and grids by the per-root counts of the
        // previous level (the union is at most their sum)
        double nf_ub = 0, fd_ub = 0;
        for (uint64_t x = td; x; x &= x - 1) {
          nf_ub += (double)h_lane[__builtin_ctzll(x)];
          fd_ub += lane_fd[__builtin_ctzll(x)];
        }
        const int64_t nf = (int64_t)std::min<double>(nf_ub, (double)rows);
        const uint64_t E_ub = (uint64_t)std::min<double>(fd_ub, (double)g.nnz);
        m.ulist_count.alloc(1);
        device_scan(BitsCountN{m.nz_sel.p, words}, BitsEmitN{m.nz_sel.p, words, m.queue.p, m.ulist_count.p},
                    (size_t)words + 1, m.scan_tmp.p, s);
        device_scan(FrontierDegreeN{m.queue.p, g.offsets.p, m.ulist_count.p}, StoreU64{m.prefix.p}, (size_t)nf + 1,
                    m.scan_tmp.p, s);
        const uint64_t nchunks = (E_ub + kEpb - 1) / kEpb;
        G5_CHECK(2 * nchunks <= m.chunk_rows.n, "g5_run_bfs_batch: chunk table too small");
        if (nchunks) {
          hipLaunchKernelGGL(k_chunk_rows<kEpb>, dim3(grid_of((int64_t)(2 * nchunks), kBlock, 65536)), dim3(kBlock),
                             0, s, m.prefix.p, nf, nchunks, m.chunk_rows.p);
          hipLaunchKernelGGL(k_ms_push, dim3(grid_of((int64_t)E_ub, kEpb, 8192)), dim3(kBlock), 0, s, m.queue.p, nf,
                             m.prefix.p, m.chunk_rows.p, g.offsets.p, g.cols.p, front, next, td);
        }
      } else {
        // over the CSR chunks holding one of the pushing roots' frontier vertices
        m.chunk_list.alloc(m.sweep_chunks);
        hipLaunchKernelGGL(k_ms_collect, dim3(grid_of((int64_t)m.sweep_chunks, kBlock, 4096)), dim3(kBlock), 0, s,
                           m.sweep_rows.p, m.sweep_chunks, m.nz_sel.p, m.chunk_list.p, d_list_count);
        hipLaunchKernelGGL(k_ms_push_list, dim3(grid_of((int64_t)m.sweep_chunks, 1, 4096)), dim3(kBlock), 0, s,
                           m.chunk_list.p, (const unsigned*)d_list_count, m.sweep_rows.p, (uint64_t)g.nnz,
                           g.offsets.p, g.cols.p, front, m.nz_sel.p, next, td);
      }
      G5_HIP(hipGetLastError());
      // probe unrolling as in the plain pull (pull_unroll)
      decltype(&k_ms_pull_lanes<4, 2>) kpl = k_ms_pull_lanes<4, 2>;
      switch (m.pull_unroll) {
        case 1: kpl = k_ms_pull_lanes<8, 2>; break;
        case 2: kpl = k_ms_pull_lanes<8, 4>; break;
        case 3: kpl = k_ms_pull_lanes<4, 4>; break;
        case 4: kpl = k_ms_pull_lanes<2, 1>; break;
        case 5: kpl = k_ms_pull_lanes<1, 1>; break;
        case 6: kpl = k_ms_pull_lanes<2, 2>; break;
        case 7: kpl = k_ms_pull_lanes<1, 2>; break;
        default: break;
      }
      hipLaunchKernelGGL(kpl, dim3(grid_of(words, kBlock / 64, m.pull_grid)), dim3(kBlock), 0, s, m.seen.p, front,
                         nz_front, next, nz_next, g.first_nbr.p, g.offsets.p, g.cols.p, rows, bu,
                         m.pull_filter == 1 ? 1 : 0, 1, cnt, lepi);
    } else if (push) {
      device_scan(UnionCount{nz_front}, UnionEmit{nz_front, m.queue.p}, (size_t)words, m.scan_tmp.p, s);
      const int64_t nf = (int64_t)nz;
      device_scan(FrontierDegree{m.queue.p, g.offsets.p, nf}, StoreU64{m.prefix.p}, (size_t)nf + 1,
                  m.scan_tmp.p, s);
      const uint64_t nchunks = (fdeg + kEpb - 1) / kEpb;
      G5_CHECK(2 * nchunks <= m.chunk_rows.n, "g5_run_bfs_batch: chunk table too small");
      if (nchunks)
        hipLaunchKernelGGL(k_chunk_rows<kEpb>, dim3(grid_of((int64_t)(2 * nchunks), kBlock, 65536)),
                           dim3(kBlock), 0, s, m.prefix.p, nf, nchunks, m.chunk_rows.p);
      ms_clear_frontier(m, l, rows, s, marked);
      // a push whose frontier edges are few next to the rows lists what it
      // touches, and only those vertices are folded
      const uint64_t touch_cap = std::min<uint64_t>(fdeg, (uint64_t)rows);
      const bool touch = (double)fdeg * kTouchBelow < (double)rows;
      if (touch) {
        m.touch.alloc(std::max<uint64_t>(touch_cap, 1));
        unsigned long long* touch_count = cnt + M_NUM;  // cleared with the counters
        G5_HIP(hipMemsetAsync(nz_next, 0, words * 8, s));
        hipLaunchKernelGGL(k_ms_push_touch, dim3(grid_of((int64_t)fdeg, kEpb, 2048)), dim3(kBlock), 0, s,
                           m.queue.p, nf, m.prefix.p, m.chunk_rows.p, g.offsets.p, g.cols.p, front, next, m.touch.p,
                           touch_count);
        hipLaunchKernelGGL(k_ms_finish_list, dim3(grid_of((int64_t)touch_cap, kBlock, 8192)), dim3(kBlock), 0, s,
                           m.seen.p, next, nz_next, g.offsets.p, m.touch.p, (const unsigned long long*)touch_count,
                           cnt, lepi, lane_stats ? 1 : 0);
      } else {
        hipLaunchKernelGGL(k_ms_push, dim3(grid_of((int64_t)fdeg, kEpb, 8192)), dim3(kBlock), 0, s, m.queue.p,
                           nf, m.prefix.p, m.chunk_rows.p, g.offsets.p, g.cols.p, front, next, ~UINT64_C(0));
      }
      G5_HIP(hipGetLastError());
      if (touch) {
        // folded above
      } else if (lane_stats)
        hipLaunchKernelGGL(k_ms_finish_lanes, dim3(grid_of(words, kBlock / 64, 8192)), dim3(kBlock), 0, s, m.seen.p,
                           next, nz_next, g.offsets.p, rows, cnt, lepi);
      else
        hipLaunchKernelGGL(k_ms_finish, dim3(grid_of(words, kBlock / 64, 8192)), dim3(kBlock), 0, s, m.seen.p,
                           next, nz_next, g.offsets.p, rows, cnt);
    } else if (sweep) {
      hipLaunchKernelGGL(k_ms_sweep_clear, dim3(grid_of((int64_t)m.sweep_chunks, kBlock, 4096)), dim3(kBlock), 0,
                         s, m.sweep_rows.p, m.sweep_chunks, g.offsets.p, next);
      hipLaunchKernelGGL(k_ms_sweep, dim3(grid_of((int64_t)m.sweep_chunks, 1, m.sweep_grid)), dim3(kBlock), 0, s,
                         m.sweep_rows.p, (uint64_t)g.nnz, g.offsets.p, g.cols.p, front, nz_front, next);
      G5_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_ms_finish, dim3(grid_of(words, kBlock / 64, 8192)), dim3(kBlock), 0, s, m.seen.p,
                         next, nz_next, g.offsets.p, rows, cnt);
    } else {
      int mode = m.pull_filter == 1 ? kFrontFilter : (m.pull_filter == 2 ? kFrontHybrid : kFrontPlain);
      const uint32_t filter_from = (uint32_t)std::min<double>((double)rows, m.filter_from * (double)rows);
      if (m.compact_below > 0 && (double)nz < m.compact_below * (double)rows) {
        // frontier words of few vertices: pack them densely
        m.pack.alloc(2 * words);
        m.compact.alloc(rows);
        device_scan(CompactCount{nz_front}, CompactEmit{nz_front, front, m.pack.p, m.compact.p}, (size_t)words,
                    m.scan_tmp.p, s);
        mode = kFrontCompact;
      }
      // probe unrolling: (short-row probes per step, long-row probes per lane and step)
      decltype(&k_ms_pull<4, 2>) kp = k_ms_pull<4, 2>;
      switch (m.pull_unroll) {
        case 1: kp = k_ms_pull<8, 2>; break;
        case 2: kp = k_ms_pull<8, 4>; break;
        case 3: kp = k_ms_pull<4, 4>; break;
        case 4: kp = k_ms_pull<2, 1>; break;
        case 5: kp = k_ms_pull<1, 1>; break;
        case 6: kp = k_ms_pull<2, 2>; break;
        case 7: kp = k_ms_pull<1, 2>; break;
        default: break;
      }
      // Tail levels: once fewer than rows / kListBelow (root, vertex) pairs
      // are unreached, a pull also marks the vertices still missing a live
      // root (one bit per vertex in m.miss, written 64 at a time), and the
      // next pull takes only those, listed by a scan of the bits (after
      // clearing its frontier array, which must hold zeros elsewhere).
      const double unreached = std::max(0.0, total_pairs - reached);
      const int64_t grid_cap =
          unreached * kLightBelow < total_pairs ? std::min<int64_t>(m.pull_grid, kLightGrid) : m.pull_grid;
      if (have_miss && mode == kFrontPlain) {
        const uint64_t ub = (uint64_t)std::min<double>((double)rows, unreached);  // list length bound
        m.ulist.alloc(std::max<uint64_t>(ub, 1));
        m.ulist_count.alloc(1);
        device_scan(BitsCountN{m.miss.p, words}, BitsEmitN{m.miss.p, words, m.ulist.p, m.ulist_count.p},
                    (size_t)words + 1, m.scan_tmp.p, s);
        ms_clear_frontier(m, l, rows, s, marked);
        G5_HIP(hipMemsetAsync(nz_next, 0, words * 8, s));
        G5_HIP(hipMemsetAsync(m.miss.p, 0, words * 8, s));
        hipLaunchKernelGGL(k_ms_pull_list, dim3(grid_of((int64_t)ub, kBlock, grid_cap)), dim3(kBlock), 0, s,
                           m.seen.p, front, next, nz_next, g.first_nbr.p, g.offsets.p, g.cols.p, m.ulist.p,
                           m.ulist_count.p, live, m.miss.p, cnt);
        list_level = true;
      } else {
        const bool mark = unreached * kListBelow < (double)rows;
        if (mark) m.miss.alloc(words);
        // the pull writes every word of the level's frontier array (clearing
        // its old words first, one store per marked vertex, so that the pull
        // could skip its zero words, measured no faster)
        hipLaunchKernelGGL(kp, dim3(grid_of(words, kBlock / 64, grid_cap)), dim3(kBlock), 0, s, m.seen.p, front,
                           nz_front, next, nz_next, g.first_nbr.p, g.offsets.p, g.cols.p, rows, live, mode,
                           m.pack.p, m.compact.p, filter_from, cnt, mark ? m.miss.p : nullptr);
        list_level = mark;
      }
    }
    G5_HIP(hipGetLastError());
    m.fr_exact[l] = 1;  // every level form leaves next non-zero exactly at the bits of nz_next
    G5_HIP(hipEventRecord(m.lev_end[l], s));
    ms_publish_wait(m, lane_stats ? kMsCounters : M_NUM, s);
    st.direction[l] = (lanes_mode && td && bu) ? 3 : (push ? 1 : (sweep ? 2 : 0));
    if (lane_stats)
      for (int r = 0; r < 64; ++r) {
        lane_fd[r] = (double)h_lane[64 + r];
        lane_vd[r] += (double)h_lane[64 + r];
      }
    have_miss = list_level;  // the next pull level can take the listed vertices only
    st.pairs[l] = (int64_t)m.h_counters[M_PAIRS];
    reached += (double)m.h_counters[M_PAIRS];
    st.num_levels = l + 1;
    live = m.h_counters[M_LIVE];
    nz = m.h_counters[M_NZ];
    m.fr_count[l] = (int64_t)nz;
    fdeg = m.h_counters[M_FDEG];
  }
  // level times, once the last level's events are complete (a level's
  // published counters can reach the host before its end event is)
  if (st.num_levels - 1 >= l0) {
    G5_HIP(hipEventSynchronize(m.lev_end[st.num_levels - 1]));
    for (int k = l0; k < st.num_levels; ++k) {
      float lms = 0;
      G5_HIP(hipEventElapsedTime(&lms, m.lev_begin[k], m.lev_end[k]));
      st.level_ms[k] = lms;
    }
  }
}

// Sets up (once per graph and tuning) the persistent path's buffers and the
// hipGraph holding the clears, the root placement and the persistent launch.
static void ms_persistent_setup(g5_ctx* ctx) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  hipStream_t s = ctx->stream;
  const int64_t rows = g.n_rows;
  const int64_t words = (rows + 63) / 64;
  const int K = m.persist_levels;
  if (m.graph_exec && m.graph_levels == K && m.graph_alpha == m.alpha && m.graph_sweep == m.sweep_below &&
      m.graph_grid == m.persist_grid && m.graph_waves == m.persist_waves && m.graph_filter == m.pull_filter &&
      m.graph_beta == m.beta)
    return;
  if (m.graph_exec) {
    (void)hipGraphExecDestroy(m.graph_exec);
    m.graph_exec = nullptr;
  }
  std::vector<uint64_t*> lev(K);
  for (int l = 0; l < K; ++l) lev[l] = ms_level(m, l, rows);
  m.p_fr.alloc(K);
  G5_HIP(hipMemcpyAsync(m.p_fr.p, lev.data(), K * sizeof(uint64_t*), hipMemcpyHostToDevice, s));
  G5_HIP(hipStreamSynchronize(s));  // lev is pageable and local
  m.p_cnt.alloc((size_t)K * M_NUM);
  m.p_ts.alloc(K);
  m.p_dir.alloc(K);
  m.p_misc.alloc(4);  // [0]: done level, [3]: active chunk count
  m.p_bar.alloc(sizeof(GridBarrier) / 8);
  m.chunk_list.alloc(m.sweep_chunks);
  if (!m.h_persist) {
    // counters [levels][M_NUM] | clock [levels] | done level + barrier (4) | roots (64)
    G5_HIP(hipHostMalloc((void**)&m.h_persist, ((size_t)G5_MAX_LEVELS * (M_NUM + 2) + 64) * sizeof(uint64_t),
                         hipHostMallocDefault));
  }
  // co-resident grid: occupancy x compute units, capped by the tuning
  int per_cu = 0, cus = 0;
  const void* kern = nullptr;
  switch (m.persist_waves) {
    case 4: kern = (const void*)k_ms_persistent<4>; break;
    case 6: kern = (const void*)k_ms_persistent<6>; break;
    default: kern = (const void*)k_ms_persistent<8>; break;
  }
  G5_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
  G5_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  G5_CHECK(per_cu >= 1, "g5_run_bfs_batch: persistent kernel does not fit a compute unit");
  int64_t grid = (int64_t)per_cu * cus;
  G5_CHECK(grid <= (int64_t)kBarGroup * kBarMaxGroups, "g5_run_bfs_batch: persistent grid beyond the barrier's groups");
  if (m.persist_grid > 0 && m.persist_grid < grid) grid = m.persist_grid;
  m.p_grid = (unsigned)grid;
  int rate_khz = 0;
  G5_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
  m.clock_khz = rate_khz > 0 ? rate_khz : 100000;

  MsPersistArgs a{};
  a.offsets = g.offsets.p;
  a.cols = g.cols.p;
  a.first_nbr = g.first_nbr.p;
  a.sweep_rows = m.sweep_rows.p;
  a.rows = rows;
  a.nnz = (uint64_t)g.nnz;
  a.sweep_chunks = m.sweep_chunks;
  a.seen = m.seen.p;
  a.fr = m.p_fr.p;
  a.nz0 = m.nz[0].p;
  a.nz1 = m.nz[1].p;
  a.cnt = (unsigned long long*)m.p_cnt.p;
  a.ts = (long long*)m.p_ts.p;
  a.dir = (int*)m.p_dir.p;
  a.done_levels = (int*)m.p_misc.p;
  a.bar = (GridBarrier*)m.p_bar.p;
  a.chunk_list = m.chunk_list.p;
  a.chunk_count = (unsigned*)(m.p_misc.p + 3);
  a.push_limit = (double)g.nnz / m.alpha;
  a.sweep_below = m.sweep_below;
  a.pull_filter = m.pull_filter;
  a.pull_unseen_frac = m.beta > 0 ? 1.0 / m.beta : 0.0;
  a.max_levels = K;
  a.timeout_ticks = (long long)m.clock_khz * 1000LL * 5;  // 5 s per barrier wait

  // The graph: clears and root placement captured from the stream, then the
  // persistent kernel as a cooperative kernel node (the runtime checks that
  // the whole grid is co-resident, which the grid barrier relies on).
  hipGraph_t graph = nullptr;
  G5_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(k_ms_fill, dim3(grid_of((int64_t)K * M_NUM, kBlock, 64)), dim3(kBlock), 0, s, m.p_cnt.p,
                     (int64_t)K * M_NUM);
  hipLaunchKernelGGL(k_ms_fill, dim3(1), dim3(kBlock), 0, s, m.p_misc.p, (int64_t)4);
  hipLaunchKernelGGL(k_ms_fill, dim3(grid_of((int64_t)(sizeof(GridBarrier) / 8), kBlock, 64)), dim3(kBlock), 0, s,
                     m.p_bar.p, (int64_t)(sizeof(GridBarrier) / 8));
  hipLaunchKernelGGL(k_ms_clear, dim3(grid_of(rows, kBlock, 4096)), dim3(kBlock), 0, s, m.seen.p, lev[0], rows);
  hipLaunchKernelGGL(k_ms_fill, dim3(grid_of(words, kBlock, 1024)), dim3(kBlock), 0, s, m.nz[0].p, words);
  hipLaunchKernelGGL(k_ms_roots, dim3(1), dim3(64), 0, s, g.old2new.p, m.roots.p, 64, rows, g.offsets.p, m.seen.p,
                     lev[0], (unsigned long long*)m.p_cnt.p);
  hipLaunchKernelGGL(k_ms_roots_nz, dim3(1), dim3(64), 0, s, g.old2new.p, m.roots.p, 64, rows, m.nz[0].p);
  const hipError_t le = hipGetLastError();
  const hipError_t ce = hipStreamEndCapture(s, &graph);
  G5_HIP(le);
  G5_HIP(ce);
  static_assert(sizeof(MsPersistArgs) <= sizeof(m.p_args_storage), "argument block too large");
  std::memcpy(m.p_args_storage, &a, sizeof(a));
  size_t nnodes = 0;
  hipError_t ge = hipGraphGetNodes(graph, nullptr, &nnodes);
  std::vector<hipGraphNode_t> deps(nnodes);
  if (ge == hipSuccess && nnodes) ge = hipGraphGetNodes(graph, deps.data(), &nnodes);
  hipGraphNode_t knode = nullptr;
  if (ge == hipSuccess) {
    hipKernelNodeParams kp{};
    kp.func = (void*)kern;
    kp.gridDim = dim3(m.p_grid);
    kp.blockDim = dim3(kBlock);
    kp.sharedMemBytes = 0;
    void* kargs[] = {(void*)m.p_args_storage};
    kp.kernelParams = kargs;
    kp.extra = nullptr;
    ge = hipGraphAddKernelNode(&knode, graph, deps.data(), nnodes, &kp);
  }
  if (ge == hipSuccess) {
    hipLaunchAttributeValue v{};
    v.cooperative = 1;
    ge = hipGraphKernelNodeSetAttribute(knode, hipLaunchAttributeCooperative, &v);
  }
  if (ge != hipSuccess) {
    (void)hipGraphDestroy(graph);
    G5_HIP(ge);
  }
  const hipError_t ie = hipGraphInstantiate(&m.graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  G5_HIP(ie);
  m.graph_levels = K;
  m.graph_alpha = m.alpha;
  m.graph_sweep = m.sweep_below;
  m.graph_grid = m.persist_grid;
  m.graph_waves = m.persist_waves;
  m.graph_filter = m.pull_filter;
  m.graph_beta = m.beta;
}

// One replay of the batch graph; levels beyond the preallocated arrays go on
// in the host loop.
static void ms_run_persistent(g5_ctx* ctx, const int64_t* roots, int nroots) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  hipStream_t s = ctx->stream;
  ms_persistent_setup(ctx);
  // the persistent kernel clears and writes the frontier words of its levels
  // itself: the per-level marks no longer describe them
  std::fill(m.fr_exact.begin(), m.fr_exact.end(), 0);
  const int K = m.persist_levels;
  int64_t* h_roots = (int64_t*)(m.h_persist + (size_t)G5_MAX_LEVELS * (M_NUM + 2));
  for (int r = 0; r < 64; ++r) h_roots[r] = r < nroots ? roots[r] : -1;
  G5_HIP(hipMemcpyAsync(m.roots.p, h_roots, 64 * sizeof(int64_t), hipMemcpyHostToDevice, s));
  G5_HIP(hipEventRecord(m.ev0, s));
  G5_HIP(hipGraphLaunch(m.graph_exec, s));
  uint64_t* h_cnt = m.h_persist;
  int64_t* h_ts = (int64_t*)(m.h_persist + (size_t)G5_MAX_LEVELS * M_NUM);
  uint64_t* h_misc = m.h_persist + (size_t)G5_MAX_LEVELS * (M_NUM + 1);
  G5_HIP(hipMemcpyAsync(h_cnt, m.p_cnt.p, (size_t)K * M_NUM * 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipMemcpyAsync(h_ts, m.p_ts.p, (size_t)K * 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipMemcpyAsync(h_misc, m.p_misc.p, 4 * 8, hipMemcpyDeviceToHost, s));
  GridBarrier* hb = (GridBarrier*)(h_misc + 4);
  G5_HIP(hipMemcpyAsync(hb, m.p_bar.p, 16, hipMemcpyDeviceToHost, s));
  std::vector<int> h_dir(K);
  G5_HIP(hipMemcpyAsync(h_dir.data(), m.p_dir.p, (size_t)K * sizeof(int), hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));

  G5_CHECK_CODE(!hb->abort, G5_ERR_HIP,
                "g5_run_bfs_batch: a grid barrier of the persistent kernel timed out (blocks not co-resident?)");
  const int done = (int)(int32_t)(h_misc[0] & 0xFFFFFFFFu);
  g5_batch_stats& st = m.stats;
  const double tick_ms = 1.0 / (double)m.clock_khz;
  double reached = (double)__builtin_popcountll(h_cnt[M_LIVE]);
  const double roots_live = reached;
  for (int l = 1; l <= done; ++l) {
    st.direction[l] = h_dir[l];
    st.pairs[l] = (int64_t)h_cnt[(size_t)l * M_NUM + M_PAIRS];
    st.level_ms[l] = (double)(h_ts[l] - h_ts[l - 1]) * tick_ms;
    reached += (double)st.pairs[l];
  }
  st.num_levels = done + 1;
  st.persistent_grid = (int32_t)m.p_grid;
  st.persistent_levels = done;
  const uint64_t* last = h_cnt + (size_t)done * M_NUM;
  if (last[M_LIVE]) {
    // out of preallocated levels: continue level by level from the host,
    // which reads the last level's nonzero bitmap from its per-level slot
    const int64_t words = (g.n_rows + 63) / 64;
    G5_HIP(hipMemcpyAsync(ms_marks(m, done, words), m.nz[done & 1].p, words * 8, hipMemcpyDeviceToDevice, s));
    G5_HIP(hipMemsetAsync(m.counters.p, 0, kMsCounters * 8, s));  // the host levels expect them clear
    ms_host_levels(ctx, done + 1, last[M_LIVE], last[M_NZ], last[M_FDEG], reached, roots_live);
  }
  (void)g;
}

void ms_run(g5_ctx* ctx, const int64_t* roots, int nroots) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  G5_CHECK(m.prepared, "g5_run_bfs_batch: call g5_batch_prepare first");
  G5_CHECK(nroots >= 1 && nroots <= 64, "g5_run_bfs_batch: 1 to 64 roots per batch");
  for (int r = 0; r < nroots; ++r)
    G5_CHECK(roots[r] >= 0 && roots[r] < g.n, "g5_run_bfs_batch: root out of range");
  hipStream_t s = ctx->stream;
  const int64_t rows = g.n_rows;
  unsigned long long* cnt = (unsigned long long*)m.counters.p;
  m.have_result = false;
  m.level_masks = false;
  m.nroots = nroots;
  m.host_roots.assign(roots, roots + nroots);
  g5_batch_stats& st = m.stats;
  st = g5_batch_stats{};
  st.num_roots = nroots;
  st.num_levels = 1;

  if (m.persistent && m.sweep_chunks) {
    ms_run_persistent(ctx, roots, nroots);
  } else {
    uint64_t* fr0 = ms_level(m, 0, rows);
    const int64_t words = (rows + 63) / 64;
    uint64_t* nz0 = ms_marks(m, 0, words);
    G5_HIP(hipMemcpyAsync(m.roots.p, roots, nroots * sizeof(int64_t), hipMemcpyHostToDevice, s));
    G5_HIP(hipEventRecord(m.ev0, s));
    G5_HIP(hipMemsetAsync(m.counters.p, 0, kMsCounters * 8, s));  // then cleared by every publish
    G5_HIP(hipMemsetAsync(m.seen.p, 0, rows * 8, s));
    const bool marked = m.fr_exact[0] != 0;
    m.fr_exact[0] = 0;
    ms_clear_frontier(m, 0, rows, s, marked);  // the previous batch's roots, or everything
    G5_HIP(hipMemsetAsync(nz0, 0, words * 8, s));
    hipLaunchKernelGGL(k_ms_roots, dim3(1), dim3(64), 0, s, g.old2new.p, m.roots.p, nroots, rows, g.offsets.p,
                       m.seen.p, fr0, cnt);
    hipLaunchKernelGGL(k_ms_roots_nz, dim3(1), dim3(64), 0, s, g.old2new.p, m.roots.p, nroots, rows, nz0);
    G5_HIP(hipGetLastError());
    m.fr_exact[0] = 1;
    m.fr_count[0] = nroots;
    ms_publish_wait(m, M_NUM, s);
    const uint64_t live = m.h_counters[M_LIVE];
    const double roots_live = (double)__builtin_popcountll(live);
    ms_host_levels(ctx, 1, live, m.h_counters[M_NZ], m.h_counters[M_FDEG], roots_live, roots_live);
  }
  G5_HIP(hipEventRecord(m.ev1, s));
  G5_HIP(hipEventSynchronize(m.ev1));
  float ms = 0;
  G5_HIP(hipEventElapsedTime(&ms, m.ev0, m.ev1));
  st.device_ms = ms;
  m.nlevels = st.num_levels - 1 > 0 ? st.num_levels - 1 : 1;  // levels with frontier words 0..L-1
  m.have_result = true;
}

static void ms_upload_level_ptrs(MsState& m, int nlev, hipStream_t s) {
  std::vector<const uint64_t*> h(nlev);
  for (int i = 0; i < nlev; ++i) h[i] = m.fr[i]->p;
  G5_HIP(hipMemcpyAsync(m.frp.p, h.data(), nlev * sizeof(void*), hipMemcpyHostToDevice, s));
}

void ms_get_levels(g5_ctx* ctx, int lane, uint8_t* host_levels) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  G5_CHECK(m.have_result, "g5_batch_get_levels: no batch result");
  G5_CHECK(lane >= 0 && lane < m.nroots, "g5_batch_get_levels: lane out of range");
  const int nlev = m.stats.num_levels;  // frontier words exist for levels 0..num_levels-1
  ms_upload_level_ptrs(m, nlev, ctx->stream);
  ctx->bfs.level_out.alloc(g.n);
  hipLaunchKernelGGL(k_ms_levels, dim3(grid_of(g.n, kBlock, 65536)), dim3(kBlock), 0, ctx->stream, g.old2new.p,
                     g.n, g.n_rows, (const uint64_t* const*)m.frp.p, nlev, lane, m.host_roots[lane],
                     ctx->bfs.level_out.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipMemcpyAsync(host_levels, ctx->bfs.level_out.p, g.n, hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(hipStreamSynchronize(ctx->stream));
}

void ms_teps_edges(g5_ctx* ctx, int64_t* out) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  G5_CHECK(m.have_result, "g5_batch_teps_edges: no batch result");
  ctx->acc.alloc(65);
  hipStream_t s = ctx->stream;
  G5_HIP(hipMemsetAsync(ctx->acc.p, 0, 65 * 8, s));
  const uint64_t all = m.nroots == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << m.nroots) - 1);
  hipLaunchKernelGGL(k_ms_teps, dim3(grid_of(g.n_rows, kBlock, 4096)), dim3(kBlock), 0, s, m.seen.p,
                     g.weight.p, g.n_rows, all, (unsigned long long*)ctx->acc.p);
  G5_HIP(hipGetLastError());
  uint64_t h[65];
  G5_HIP(hipMemcpyAsync(h, ctx->acc.p, 65 * 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  for (int r = 0; r < m.nroots; ++r) out[r] = (int64_t)((h[r] + h[64]) / 2);
}

void ms_model_bytes(g5_ctx* ctx, double* per_level, double* total) {
  Graph& g = ctx->graph;
  MsState& m = ctx->ms;
  G5_CHECK(m.have_result, "g5_batch_model_bytes: no batch result");
  hipStream_t s = ctx->stream;
  const int nlev = m.stats.num_levels;
  ms_upload_level_ptrs(m, nlev, s);
  ctx->acc.alloc(65);
  const uint64_t lanes = m.nroots == 64 ? ~UINT64_C(0) : ((UINT64_C(1) << m.nroots) - 1);
  // The batch's output is fr[l] (written at every level, inside the 24 n
  // term); a root's uint8 levels are unpacked from it on request, outside
  // the batch, so no per-(root, vertex) output byte is counted.
  double tot = 0;
  for (int l = 0; l < G5_MAX_LEVELS; ++l)
    if (per_level) per_level[l] = 0;
  for (int l = 1; l < nlev; ++l) {
    G5_HIP(hipMemsetAsync(ctx->acc.p, 0, 8, s));
    const uint64_t pull = m.level_masks ? m.level_pull[l] & lanes : lanes;
    const uint64_t push = m.level_masks ? m.level_push[l] & lanes : 0;
    hipLaunchKernelGGL(k_ms_model, dim3(grid_of(g.n_rows, kBlock, 8192)), dim3(kBlock), 0, s,
                       (const uint64_t* const*)m.frp.p, l, pull, push, g.offsets.p, g.cols.p, g.n_rows,
                       (unsigned long long*)ctx->acc.p);
    G5_HIP(hipGetLastError());
    uint64_t h = 0;
    G5_HIP(hipMemcpyAsync(&h, ctx->acc.p, 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    const double b = 24.0 * (double)g.n + (double)h;
    if (per_level) per_level[l] = b;
    tot += b;
  }
  if (total) *total = tot;
}

}  // namespace g5
