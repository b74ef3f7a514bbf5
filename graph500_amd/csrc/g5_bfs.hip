// g5_bfs.hip -- single-root direction-optimizing BFS on gfx950.
//
// Reference behaviour restated: BfsBase::run_bfs (mpi/bfs.hpp:2651-2940):
// level-synchronous search that starts top-down, switches to bottom-up when
// the frontier grows beyond n_active/2000 and back to top-down (for good) when
// fewer than n_active/8 vertices remain unvisited (mpi/bfs.hpp:2799-2829,
// constants mpi/parameters.h:87-89).
//
// Kernels (vertex ids are degree-ordered "new" ids, rows [0, n_rows)):
//   top-down   k_td_expand: the frontier's rows are concatenated into one
//              virtual edge list (exclusive scan of frontier degrees); each
//              workgroup takes kEpb-edge chunks, stages the prefix/row-base/
//              source of the frontier vertices overlapping its chunk in LDS and
//              binary-searches there once per lane, so hub rows are split
//              across the whole grid.  Small levels test-and-set the visited
//              bitmap and append winners through LDS (one global atomic per
//              chunk); hub levels use no atomics: the visited bitmap is a
//              read-only snapshot, new targets get level/parent with benign
//              races, and one pass over the level array gives the frontier.
//              While the search grows, a hub level first settles its
//              highest-degree targets bottom-up (k_td_head) and expands only
//              the row suffixes beyond them (k_td_expand_segs).
//              After the bottom-up phase, a top-down level whose frontier is
//              large against the unvisited vertices is found by the bottom-up
//              kernel (same vertices, far fewer probes); consecutive small
//              top-down levels run fused in one single-workgroup launch
//              (k_td_fused) that applies the switch rule itself.
//   host loop  one round trip per level: a one-lane kernel publishes the
//              level's counters to coherent host memory, each value tagged
//              with the level's sequence number, where the host reads them
//              for the alpha/beta rule.
//   queues     a top-down frontier that comes from a bitmap is rebuilt by an
//              exclusive scan of per-word popcounts (ascending vertex order).
//   bottom-up  k_bu: one wave per two 64-vertex bitmap words, one lane per
//              vertex of each (both words' first probes in flight together).
//              Unvisited lanes first probe their first (highest-degree)
//              neighbour from the coalesced first_nbr array; misses with a
//              short row scan it lane-sequentially, long rows are scanned by
//              the whole wave 64 entries at a time with a ballot for the first
//              hit.  The visited snapshot of the previous level is read-only;
//              the next bitmap word is written once per wave from a ballot.
#include "g5_ctx.hpp"
#include "g5_frontier.hpp"
#include "g5_prims.hpp"

#include <algorithm>

namespace g5 {

constexpr uint64_t kTdSmall = 1 << 20;  // top-down levels up to this many edges use the atomic path
constexpr int kBinShift = 19;           // binned top-down: at least 2^19 targets per bucket
constexpr int kShortRow = 32;    // bottom-up: lane-sequential scan up to this length
constexpr uint32_t kNoNbr = 0xFFFFFFFFu;

// counters: vertices found; the root (new id); head words (k_td_head_size);
// frontier edges of the next level, summed by a small top-down level
enum { C_FOUND = 0, C_AUX = 1, C_HEAD = 2, C_NEXT_E = 3, C_NUM = 4 };


__global__ void __launch_bounds__(kBlock) k_bfs_init(uint8_t* __restrict__ level, uint64_t* __restrict__ vis,
                                                     int64_t rows, int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += stride) level[v] = 0xFF;
  for (int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x; w < words; w += stride) {
    // bits past the last row read as visited so no sweep ever picks them
    const int64_t first = w * 64;
    uint64_t x = 0;
    if (first + 64 > rows) x = (rows <= first) ? ~UINT64_C(0) : (~UINT64_C(0) << (rows - first));
    vis[w] = x;
  }
}

__global__ void k_bfs_root(const uint32_t* __restrict__ old2new, int64_t root_orig, int64_t rows,
                           uint8_t* __restrict__ level, uint32_t* __restrict__ parent,
                           uint64_t* __restrict__ vis, uint32_t* __restrict__ queue,
                           unsigned long long* __restrict__ counters) {
  const uint32_t r = old2new[root_orig];
  // the first level's counters start at zero; k_publish clears them after
  // every later level
  for (int i = 0; i < C_NUM; ++i) counters[i] = 0;
  if ((int64_t)r < rows) {
    level[r] = 0;
    parent[r] = r;
    vis[r >> 6] |= UINT64_C(1) << (r & 63);
    queue[0] = r;
  }
}

// Top-down expansion of one level.  The frontier's rows form one virtual edge
// list (prefix = exclusive scan of frontier degrees, E = prefix[nf]); every
// workgroup takes kEpb-edge chunks and each lane walks kEpt consecutive edges.
//   kBulk = false (small levels): test-and-set on the visited bitmap itself
//           (atomicOr after a plain-load filter); winners are staged in LDS
//           and appended to 'next' with one global atomic per chunk, which
//           also counts them in counters[C_FOUND]; their degree sum (the
//           next level's frontier edges) goes to counters[C_NEXT_E].
//   kBulk = true  (hub levels): no atomics: a target gets level and parent
//           unless its level is already set (racing writers store the same
//           level and a valid parent); k_td_finish then takes the exact new
//           set from the level array.  The visited-bitmap snapshot is probed
//           first only when probe_vis is set (levels after the bottom-up
//           phase, where it drops most targets); while the search grows it
//           holds few vertices and the level test alone decides.
// Rows below the returned id were settled by k_td_head (none without a head).
__device__ __forceinline__ uint32_t head_bound(const unsigned long long* head_words) {
  if (!head_words) return 0;
  const unsigned long long v = *head_words * 64;
  return v < UINT32_MAX ? (uint32_t)v : UINT32_MAX;
}

__device__ __forceinline__ void bulk_claims(uint64_t e0, uint64_t e1, int cnt, const ChunkRows& cr,
                                            const uint32_t* __restrict__ cols, const uint64_t* vis,
                                            uint8_t* level, uint32_t* __restrict__ parent, uint32_t lvl,
                                            int probe_vis, uint32_t head) {
  // every probe of the lane's edges in flight before any claim
  uint32_t w[kEpt], u[kEpt];
  bool ok[kEpt];
  gather_strided(e0, e1, cnt, cr, cols, w, u, ok);
  if (probe_vis) {
    // most targets are visited already: the bitmap (1/8 of the level
    // array's footprint) drops them before any level probe
    uint64_t vw[kEpt];
#pragma unroll
    for (int j = 0; j < kEpt; ++j) vw[j] = ok[j] ? vis[w[j] >> 6] : 0;
#pragma unroll
    for (int j = 0; j < kEpt; ++j) ok[j] = ok[j] && !((vw[j] >> (w[j] & 63)) & 1);
  }
  // targets below 'head' were settled by k_td_head
#pragma unroll
  for (int j = 0; j < kEpt; ++j) ok[j] = ok[j] && w[j] >= head;
  uint8_t lv[kEpt];
#pragma unroll
  for (int j = 0; j < kEpt; ++j) lv[j] = ok[j] ? level[w[j]] : 0;
#pragma unroll
  for (int j = 0; j < kEpt; ++j)
    if (ok[j] && lv[j] == 0xFF) {
      level[w[j]] = (uint8_t)lvl;
      parent[w[j]] = u[j];
    }
}

template <bool kBulk>
__global__ void __launch_bounds__(kBlock) k_td_expand(
    const uint32_t* __restrict__ queue, int64_t nf, const uint64_t* __restrict__ prefix,
    const uint32_t* __restrict__ chunk_rows, const uint64_t* __restrict__ offsets,
    const uint32_t* __restrict__ cols, uint64_t* vis, uint8_t* level, uint32_t* __restrict__ parent,
    uint32_t* __restrict__ next, unsigned long long* __restrict__ counters, uint32_t lvl, int probe_vis,
    const unsigned long long* __restrict__ head_words) {
  __shared__ ChunkRows cr;
  __shared__ uint32_t s_q[kBulk ? 1 : kEpb];
  __shared__ uint32_t s_qn;
  __shared__ uint64_t s_qbase;
  const uint64_t E = prefix[nf];
  // both forms are launched before the host knows E; the other one exits
  if (kBulk ? E <= kTdSmall : E > kTdSmall) return;
  const uint32_t head = head_bound(head_words);
  uint64_t next_e = 0;  // degree sum of this lane's winners (small levels)
  for (uint64_t chunk = blockIdx.x; chunk * kEpb < E; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kEpb;
    const uint64_t e1 = (e0 + kEpb < E) ? e0 + kEpb : E;
    if (!kBulk && threadIdx.x == 0) s_qn = 0;
    const int cnt = stage_chunk(chunk, chunk_rows, queue, prefix, offsets, cr);
    const auto visit = [&](uint32_t w, uint32_t u) {
      const uint64_t bit = UINT64_C(1) << (w & 63);
      if (vis[w >> 6] & bit) return;
      const uint64_t old = atomicOr((unsigned long long*)(vis + (w >> 6)), (unsigned long long)bit);
      if (old & bit) return;
      level[w] = (uint8_t)lvl;
      parent[w] = u;
      s_q[atomicAdd(&s_qn, 1u)] = w;
      next_e += offsets[w + 1] - offsets[w];
    };
    if (kBulk) {
      bulk_claims(e0, e1, cnt, cr, cols, vis, level, parent, lvl, probe_vis, head);
    } else {
      walk_strided(e0, e1, cnt, cr, cols, visit);
    }
    if (!kBulk) {
      __syncthreads();
      if (threadIdx.x == 0) s_qbase = s_qn ? atomicAdd(counters + C_FOUND, (unsigned long long)s_qn) : 0;
      __syncthreads();
      for (uint32_t j = threadIdx.x; j < s_qn; j += kBlock) next[s_qbase + j] = s_q[j];
    }
    __syncthreads();
  }
  // the next level's frontier edges, so that the host can skip the launches
  // of the other form there
  if (!kBulk) block_add(next_e, counters + C_NEXT_E);
}

// --- piece-ordered hub levels ---------------------------------------------
// A hub level's frontier rows hold up to billions of edges whose targets,
// sorted within each row, spread over the whole level array, so taking the
// rows one after another keeps the grid's probes scattered over all of it.
// Instead every row is cut into P pieces by position (piece k of a row of
// degree d holds its positions [k*d/P, (k+1)*d/P)) and pieces are taken piece
// index first.  The rows of a Kronecker graph share nearly one neighbour
// distribution (a target bit is 0 with probability A/(A+B) = 0.75 or
// C/(C+D) = 0.79 for the two values of the row bit under the 0.57/0.19/0.19
// initiator), so the k-th pieces of all rows cover about the same target
// range and the level and parent bytes the grid touches at a time stay
// cache-sized.  One scan compacts the non-empty pieces into a segment list
// (the scanned value carries the piece count above bit 40 and its edges
// below) that the chunked bulk expansion walks like frontier rows.
constexpr int kSegCountShift = 40;
constexpr uint64_t kSegEdgeMask = (UINT64_C(1) << kSegCountShift) - 1;
constexpr int64_t kSegMax = (INT64_C(1) << 24) - 1;  // segments of one level (count field width)

struct PieceLen {
  const uint32_t* queue;
  const uint64_t* offsets;
  const uint64_t* prefix;
  int64_t nf;
  int64_t pieces;
  // piece s = k * nf + i: frontier row i's piece k -> source, first cols index, length
  __device__ void piece(size_t s, uint32_t& u, uint64_t& a, uint64_t& len) const {
    const int64_t k = (int64_t)s / nf, i = (int64_t)s - k * nf;
    u = queue[i];
    const uint64_t o = offsets[u], d = offsets[u + 1] - o;
    a = o + (uint64_t)k * d / (uint64_t)pieces;
    len = o + (uint64_t)(k + 1) * d / (uint64_t)pieces - a;
  }
  __device__ uint64_t operator()(size_t s) const {
    // a small level (E <= kTdSmall) makes no segments
    if ((int64_t)s >= nf * pieces || prefix[nf] <= kTdSmall) return 0;
    uint32_t u;
    uint64_t a, len;
    piece(s, u, a, len);
    return len ? (UINT64_C(1) << kSegCountShift) | len : 0;
  }
};
// Exact target blocks (block = 2^shift consecutive target ids): the first
// hub_rows rows (the highest-degree ones in degree order) carry the cols
// index of their first entry in every block, relative to the row start,
// bnd[v * (nb + 1) + b] for b in [0, nb] (bnd[.. + nb] = degree).  Piece
// s = k * nf + i is row i's block k for a hub row and, for k = nb, the whole
// row of any other frontier vertex (whose edges are few at a hub level).
struct BlockLen {
  const uint32_t* queue;
  const uint64_t* offsets;
  const uint64_t* prefix;
  int64_t nf;
  int64_t pieces;  // nb + 1
  const uint32_t* bnd;
  int64_t hub_rows;
  __device__ void piece(size_t s, uint32_t& u, uint64_t& a, uint64_t& len) const {
    const int64_t k = (int64_t)s / nf, i = (int64_t)s - k * nf;
    u = queue[i];
    const uint64_t o = offsets[u];
    a = o;
    len = 0;
    if ((int64_t)u < hub_rows) {
      if (k + 1 < pieces) {
        const uint32_t* r = bnd + (int64_t)u * pieces + k;
        a = o + r[0];
        len = r[1] - r[0];
      }
    } else if (k + 1 == pieces) {
      len = offsets[u + 1] - o;
    }
  }
  __device__ uint64_t operator()(size_t s) const {
    if ((int64_t)s >= nf * pieces || prefix[nf] <= kTdSmall) return 0;
    uint32_t u;
    uint64_t a, len;
    piece(s, u, a, len);
    return len ? (UINT64_C(1) << kSegCountShift) | len : 0;
  }
};

// bnd of the block index (one lane per row and block; the lanes of a wave
// share a row, so the first steps of their binary searches coincide)
__global__ void __launch_bounds__(kBlock) k_hub_bounds(const uint64_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ cols, int64_t hub_rows,
                                                       int64_t pieces, int shift, uint32_t* __restrict__ bnd) {
  const int64_t total = hub_rows * pieces;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
    const int64_t v = t / pieces, b = t - v * pieces;
    const uint64_t key = (uint64_t)b << shift;
    uint64_t lo = offsets[v], hi = offsets[v + 1];
    const uint64_t o = lo;
    while (lo < hi) {  // first entry >= key
      const uint64_t mid = (lo + hi) >> 1;
      if (cols[mid] < key) lo = mid + 1; else hi = mid;
    }
    bnd[t] = (uint32_t)(lo - o);
  }
}

// With a head (k_td_head), the edges a frontier row still has to expand are
// its suffix of targets >= the head bound (rows are sorted): one segment per
// row, whose start k_suffix_start finds by binary search (one lane per row),
// so the expansion no longer reads the columns of edges into the head.
__global__ void __launch_bounds__(kBlock) k_suffix_start(const uint32_t* __restrict__ queue, int64_t nf,
                                                         const uint64_t* __restrict__ prefix,
                                                         const uint64_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ cols,
                                                         const unsigned long long* __restrict__ head_words,
                                                         uint64_t* __restrict__ start) {
  if (prefix[nf] <= kTdSmall) return;
  const uint32_t h = head_bound(head_words);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nf; i += (int64_t)gridDim.x * kBlock) {
    const uint32_t u = queue[i];
    uint64_t lo = offsets[u], hi = offsets[u + 1];
    while (lo < hi) {  // first entry >= h
      const uint64_t mid = (lo + hi) >> 1;
      if (cols[mid] < h) lo = mid + 1; else hi = mid;
    }
    start[i] = lo;
  }
}

struct SuffixLen {
  const uint32_t* queue;
  const uint64_t* offsets;
  const uint64_t* prefix;
  int64_t nf;
  int64_t pieces;         // 1
  const uint64_t* start;  // k_suffix_start
  __device__ void piece(size_t s, uint32_t& u, uint64_t& a, uint64_t& len) const {
    u = queue[s];
    a = start[s];
    len = offsets[u + 1] - a;
  }
  __device__ uint64_t operator()(size_t s) const {
    if ((int64_t)s >= nf || prefix[nf] <= kTdSmall) return 0;
    uint32_t u;
    uint64_t a, len;
    piece(s, u, a, len);
    return len ? (UINT64_C(1) << kSegCountShift) | len : 0;
  }
};

template <class Len>
struct PieceEmit {
  Len in;
  uint64_t* seg_pre;   // first virtual edge of each segment; seg_pre[count] = E
  uint64_t* seg_base;  // cols index - seg_pre
  uint32_t* seg_src;
  uint64_t* seg_count;
  __device__ void operator()(size_t s, uint64_t pre) const {
    const uint64_t pos = pre >> kSegCountShift, ep = pre & kSegEdgeMask;
    if ((int64_t)s == in.nf * in.pieces) {  // the extra last element carries the totals
      seg_count[0] = pos;
      seg_pre[pos] = ep;
      return;
    }
    if (in.prefix[in.nf] <= kTdSmall) return;
    uint32_t u;
    uint64_t a, len;
    in.piece(s, u, a, len);
    if (!len) return;
    seg_pre[pos] = ep;
    seg_base[pos] = a - ep;
    seg_src[pos] = u;
  }
};

// k_chunk_rows over the segment list, whose length is known on the device only.
__global__ void __launch_bounds__(kBlock) k_chunk_segs(const uint64_t* __restrict__ seg_pre,
                                                       const uint64_t* __restrict__ seg_count, uint64_t nchunks,
                                                       uint32_t* __restrict__ rows_out) {
  const int64_t ns = (int64_t)seg_count[0];
  if (ns == 0) return;
  const uint64_t E = seg_pre[ns];
  const uint64_t nc = (E + kEpb - 1) / kEpb < nchunks ? (E + kEpb - 1) / kEpb : nchunks;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < 2 * nc; t += (uint64_t)gridDim.x * kBlock) {
    const uint64_t c = t >> 1;
    uint64_t key = c * kEpb;
    if (t & 1) key = ((key + kEpb < E) ? key + kEpb : E) - 1;
    int64_t lo = 0, hi = ns - 1;  // last segment with seg_pre <= key
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (seg_pre[mid] <= key) lo = mid; else hi = mid - 1;
    }
    rows_out[t] = (uint32_t)lo;
  }
}

// The bulk expansion over the segment list: a bulk level (E = prefix[nf] >
// kTdSmall) walks the seg_pre[count] edges of its segments.
__global__ void __launch_bounds__(kBlock) k_td_expand_segs(
    const uint64_t* __restrict__ prefix, int64_t nf, const uint64_t* __restrict__ seg_pre,
    const uint64_t* __restrict__ seg_base, const uint32_t* __restrict__ seg_src,
    const uint64_t* __restrict__ seg_count, const uint32_t* __restrict__ chunk_rows,
    const uint32_t* __restrict__ cols, const uint64_t* vis, uint8_t* level, uint32_t* __restrict__ parent,
    uint32_t lvl, const unsigned long long* __restrict__ head_words) {
  __shared__ ChunkRows cr;
  if (prefix[nf] <= kTdSmall) return;
  const uint64_t E = seg_pre[*seg_count];
  const uint32_t head = head_bound(head_words);
  for (uint64_t chunk = blockIdx.x; chunk * kEpb < E; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kEpb;
    const uint64_t e1 = (e0 + kEpb < E) ? e0 + kEpb : E;
    // every segment is non-empty, so at most kEpb + 1 of them overlap a chunk
    const int64_t i0 = chunk_rows[2 * chunk];
    int cnt = (int)((int64_t)chunk_rows[2 * chunk + 1] - i0 + 1);
    cnt = cnt > kEpb + 1 ? kEpb + 1 : cnt;
    for (int j = threadIdx.x; j < cnt; j += kBlock) {
      cr.pre[j] = seg_pre[i0 + j];
      cr.base[j] = seg_base[i0 + j];
      cr.src[j] = seg_src[i0 + j];
    }
    __syncthreads();
    bulk_claims(e0, e1, cnt, cr, cols, vis, level, parent, lvl, 0, head);
    __syncthreads();
  }
}

// Row part of the bottom-up probe: the first entry of [offsets[v] + 1,
// offsets[v + 1]) (the first neighbour is probed apart, from the coalesced
// first_nbr array) whose bit is set in bm.  Every lane of the wave calls it;
// lanes with active=false only help.  Rows up to kShortRow are scanned by
// their lane, longer rows by the whole wave 64 entries at a time with a ballot
// for the first hit.
__device__ __forceinline__ bool row_marked_nbr(bool active, uint32_t v, const uint64_t* __restrict__ bm,
                                               const uint64_t* __restrict__ offsets,
                                               const uint32_t* __restrict__ cols, uint32_t& p) {
  const int lane = threadIdx.x & 63;
  bool found = false, need = false;
  uint64_t rb = 0, re = 0;
  if (active) {
    rb = offsets[v] + 1;
    re = offsets[v + 1];
    need = rb < re;
  }
  if (need && re - rb <= (uint64_t)kShortRow) {
    // four independent probes per step (entries past the row repeat the
    // first); any marked neighbour is a valid parent, the first is taken
    for (uint64_t k = rb; k < re && !found; k += 4) {
      const uint32_t c0 = cols[k];
      const uint32_t c1 = k + 1 < re ? cols[k + 1] : c0;
      const uint32_t c2 = k + 2 < re ? cols[k + 2] : c0;
      const uint32_t c3 = k + 3 < re ? cols[k + 3] : c0;
      const bool h0 = bit_set(bm, c0), h1 = bit_set(bm, c1), h2 = bit_set(bm, c2), h3 = bit_set(bm, c3);
      if (h0 | h1 | h2 | h3) {
        found = true;
        p = h0 ? c0 : (h1 ? c1 : (h2 ? c2 : c3));
      }
    }
    need = false;
  }
  uint64_t lm = __ballot(need);
  while (lm) {
    const int src = __ffsll((unsigned long long)lm) - 1;
    lm &= lm - 1;
    const uint64_t b = __shfl(rb, src, 64), e = __shfl(re, src, 64);
    // 128 entries per step, two independent probes per lane
    for (uint64_t base = b; base < e; base += 128) {
      const uint64_t k = base + lane;
      uint32_t c0 = 0, c1 = 0;
      bool h0 = false, h1 = false;
      if (k < e) c0 = cols[k];
      if (k + 64 < e) c1 = cols[k + 64];
      if (k < e) h0 = bit_set(bm, c0);
      if (k + 64 < e) h1 = bit_set(bm, c1);
      const uint64_t m0 = __ballot(h0), m1 = __ballot(h1);
      if (m0 | m1) {
        const uint32_t hp = m0 ? __shfl(c0, __ffsll((unsigned long long)m0) - 1, 64)
                               : __shfl(c1, __ffsll((unsigned long long)m1) - 1, 64);
        if (lane == src) {
          found = true;
          p = hp;
        }
        break;
      }
    }
  }
  return found;
}

// First neighbour of v, in ascending (highest degree first) order, whose bit
// is set in bm: the bottom-up probe (first_nbr, then the row).  Every lane of
// the wave calls it; lanes with active=false only help.
__device__ __forceinline__ bool first_marked_nbr(bool active, uint32_t v, const uint64_t* __restrict__ bm,
                                                 const uint32_t* __restrict__ first_nbr,
                                                 const uint64_t* __restrict__ offsets,
                                                 const uint32_t* __restrict__ cols, uint32_t& p) {
  const uint32_t f = active ? first_nbr[v] : kNoNbr;
  const bool hit = f != kNoNbr && bit_set(bm, f);
  if (hit) p = f;
  const bool more = row_marked_nbr(f != kNoNbr && !hit, v, bm, offsets, cols, p);
  return hit || more;
}

// Head of a growing bulk top-down level: the rows [0, 64 * head_words) (the
// highest-degree vertices in degree order) are settled bottom-up before the
// expansion, which then skips every edge into them.  An unvisited vertex
// whose neighbour is visited is reached at this level, because a neighbour
// from an older level would have reached it earlier, so the visited bitmap
// stands for the frontier and the vertex gets the same level it would get
// top-down, with a valid parent.  The hubs' rows are long, but most of them
// meet a visited neighbour within their first (highest-degree) entries,
// whereas top-down every frontier row probes them again and again.
__global__ void __launch_bounds__(kBlock) k_td_head(const uint64_t* __restrict__ vis,
                                                    const uint32_t* __restrict__ first_nbr,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ cols, uint8_t* __restrict__ level,
                                                    uint32_t* __restrict__ parent,
                                                    const unsigned long long* __restrict__ head_words_p, uint32_t lvl) {
  const int64_t head_words = (int64_t)*head_words_p;  // 0 at a small level
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t w = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < head_words; w += nwaves) {
    const uint64_t cur = vis[w];
    if (cur == ~UINT64_C(0)) continue;
    const uint32_t v = (uint32_t)(w * 64 + lane);
    uint32_t p = 0;
    if (first_marked_nbr(!((cur >> lane) & 1), v, vis, first_nbr, offsets, cols, p)) {
      level[v] = (uint8_t)lvl;
      parent[v] = p;
    }
  }
}

// How far the head reaches, from the level's frontier edge count E: a row
// of degree d belongs to it when d >= factor * (nnz / E)^2.  Settling such a
// vertex bottom-up takes about nnz / E probes (a random neighbour is a
// frontier vertex with probability about E / nnz), while top-down its
// level byte is probed about d * E / nnz times; the two balance at
// d = (nnz / E)^2.  Rows are in descending degree order, so the head is the
// prefix of rows above the threshold (found by binary search), in bitmap
// words.  A small level (E <= kTdSmall) gets none.
__global__ void k_td_head_size(const uint64_t* __restrict__ prefix, int64_t nf,
                               const uint64_t* __restrict__ offsets, int64_t rows, int64_t words, double nnz,
                               double factor, unsigned long long* __restrict__ head_words) {
  const uint64_t E = prefix[nf];
  int64_t h = 0;
  if (E > kTdSmall) {
    const double r = nnz / (double)E;
    const double thr = factor * r * r;
    int64_t lo = 0, hi = rows;  // first row with degree < thr
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((double)(offsets[mid + 1] - offsets[mid]) >= thr) lo = mid + 1; else hi = mid;
    }
    h = (lo + 63) / 64;
    if (h > words) h = words;
  }
  *head_words = (unsigned long long)h;
}

// Binned top-down for hub levels (hundreds of millions of frontier edges):
// the edges' (target, source) pairs are first written out grouped by
// target range (bucket = target >> shift, 2^shift targets, whose level and
// parent bytes fit in L2), then applied bucket after bucket, so the
// random level probes of the plain bulk path become L2-local.  Two passes
// over the frontier's rows (count, then scatter with per-block ranges
// reserved by one atomic per bucket) and one over the pairs.
constexpr int kBinMaxBuckets = 2048;

__global__ void __launch_bounds__(kBlock) k_td_bin_count(const uint32_t* __restrict__ queue, int64_t nf,
                                                         const uint64_t* __restrict__ prefix,
                                                         const uint32_t* __restrict__ chunk_rows,
                                                         const uint64_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ cols, int shift, int nb,
                                                         unsigned long long* __restrict__ bucket_cnt) {
  __shared__ ChunkRows cr;
  __shared__ unsigned s_hist[kBinMaxBuckets];
  for (int i = threadIdx.x; i < nb; i += kBlock) s_hist[i] = 0;
  __syncthreads();
  const uint64_t E = prefix[nf];
  for (uint64_t chunk = blockIdx.x; chunk * kEpb < E; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kEpb;
    const uint64_t e1 = (e0 + kEpb < E) ? e0 + kEpb : E;
    const int cnt = stage_chunk(chunk, chunk_rows, queue, prefix, offsets, cr);
    walk_strided(e0, e1, cnt, cr, cols, [&](uint32_t w, uint32_t) { atomicAdd(&s_hist[w >> shift], 1u); });
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nb; i += kBlock)
    if (s_hist[i]) atomicAdd(bucket_cnt + i, (unsigned long long)s_hist[i]);
}

// bucket_cnt -> exclusive offsets in bucket_cur (one block; nb <= kBinMaxBuckets)
__global__ void __launch_bounds__(kBlock) k_td_bin_offsets(const unsigned long long* __restrict__ bucket_cnt,
                                                           int nb, unsigned long long* __restrict__ bucket_cur) {
  if (threadIdx.x == 0) {
    unsigned long long run = 0;
    for (int i = 0; i < nb; ++i) {
      bucket_cur[i] = run;
      run += bucket_cnt[i];
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_td_bin_scatter(const uint32_t* __restrict__ queue, int64_t nf,
                                                           const uint64_t* __restrict__ prefix,
                                                           const uint32_t* __restrict__ chunk_rows,
                                                           const uint64_t* __restrict__ offsets,
                                                           const uint32_t* __restrict__ cols, int shift, int nb,
                                                           unsigned long long* __restrict__ bucket_cur,
                                                           uint64_t* __restrict__ pairs) {
  __shared__ ChunkRows cr;
  __shared__ unsigned long long s_cur[kBinMaxBuckets];
  __shared__ unsigned s_hist[kBinMaxBuckets];
  const uint64_t E = prefix[nf];
  for (uint64_t chunk = blockIdx.x; chunk * kEpb < E; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kEpb;
    const uint64_t e1 = (e0 + kEpb < E) ? e0 + kEpb : E;
    for (int i = threadIdx.x; i < nb; i += kBlock) s_hist[i] = 0;
    const int cnt = stage_chunk(chunk, chunk_rows, queue, prefix, offsets, cr);  // ends with a barrier
    walk_strided(e0, e1, cnt, cr, cols, [&](uint32_t w, uint32_t) { atomicAdd(&s_hist[w >> shift], 1u); });
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kBlock)
      s_cur[i] = s_hist[i] ? atomicAdd(bucket_cur + i, (unsigned long long)s_hist[i]) : 0;
    __syncthreads();
    walk_strided(e0, e1, cnt, cr, cols, [&](uint32_t w, uint32_t u) {
      const unsigned long long pos = atomicAdd(&s_cur[w >> shift], 1ull);
      pairs[pos] = ((uint64_t)w << 32) | u;
    });
    __syncthreads();
  }
}

// Applies the pairs in order.  The grid strides over them together, so the
// pairs in flight at any moment span a window of grid * kBlock pairs, i.e.
// one or two target ranges, whose level and parent lines stay in L2.
__global__ void __launch_bounds__(kBlock) k_td_bin_apply(const uint64_t* __restrict__ pairs, uint64_t E,
                                                         uint8_t* level, uint32_t* __restrict__ parent,
                                                         uint32_t lvl) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < E; i += stride) {
    const uint64_t x = pairs[i];
    const uint32_t w = (uint32_t)(x >> 32);
    if (level[w] != 0xFF) continue;
    level[w] = (uint8_t)lvl;
    parent[w] = (uint32_t)x;
  }
}

// After a bulk top-down level: the vertices whose level became 'lvl' join the
// visited bitmap and form the frontier bitmap; their number goes to counters.
// A wave takes 16 bitmap words (1024 vertices) per step: lane i reads 16
// level bytes with one 16-byte load (1 KiB per wave instruction), and lanes
// 4k..4k+3 assemble word k with two shuffles.
__device__ __forceinline__ uint32_t bytes_equal4(uint32_t x, uint32_t pat) {
  const uint32_t y = x ^ pat;
  uint32_t r = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) r |= (((y >> (8 * b)) & 0xFFu) == 0 ? 1u : 0u) << b;
  return r;
}

__global__ void __launch_bounds__(kBlock) k_td_finish(const uint8_t* __restrict__ level, int64_t rows,
                                                      int64_t words, uint32_t lvl, uint64_t* __restrict__ vis,
                                                      uint64_t* __restrict__ fbits,
                                                      unsigned long long* __restrict__ counters,
                                                      const uint64_t* __restrict__ prefix, int64_t nf) {
  if (prefix[nf] <= kTdSmall) return;  // a small level: k_td_expand<false> kept the queue exact
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
  const int64_t groups = (words + 15) / 16;
  const uint32_t pat = (lvl & 0xFFu) * 0x01010101u;
  uint64_t total = 0;
  for (int64_t gi = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); gi < groups; gi += nwaves) {
    const int64_t v0 = gi * 1024 + lane * 16;
    uint32_t m16 = 0;
    if (v0 + 16 <= rows) {
      const uint4 q = *reinterpret_cast<const uint4*>(level + v0);  // v0 is a multiple of 16
      m16 = bytes_equal4(q.x, pat) | (bytes_equal4(q.y, pat) << 4) | (bytes_equal4(q.z, pat) << 8) |
            (bytes_equal4(q.w, pat) << 12);
    } else {
      for (int j = 0; j < 16; ++j)
        if (v0 + j < rows && level[v0 + j] == (uint8_t)lvl) m16 |= 1u << j;
    }
    uint64_t m = (uint64_t)m16 << (16 * (lane & 3));
    m |= __shfl_xor(m, 1, 64);
    m |= __shfl_xor(m, 2, 64);
    const int64_t w = gi * 16 + (lane >> 2);
    if ((lane & 3) == 0 && w < words) {
      vis[w] |= m;
      fbits[w] = m;
      total += __popcll(m);
    }
  }
  block_add(total, counters + C_FOUND);
}

// Bottom-up level.  A wave reads `span` bitmap words at a time (one
// coalesced load, one word per lane, span a power of two from 2 to 64),
// copies the words whose vertices are all visited to the next bitmap (one
// coalesced store) and walks the others two at a time, one lane per vertex
// of each, so both words' first-neighbour probes are in flight together.
// The host sizes span so that the grid still has a step per wave or more;
// late levels, where almost every word is full, then cost little more than
// a pass over the bitmap.
__global__ void __launch_bounds__(kBlock) k_bu(
    const uint64_t* __restrict__ vis, uint64_t* __restrict__ vis_next,
    const uint32_t* __restrict__ first_nbr, const uint64_t* __restrict__ offsets,
    const uint32_t* __restrict__ cols, uint8_t* __restrict__ level, uint32_t* __restrict__ parent,
    int64_t words, int span, uint32_t lvl, unsigned long long* __restrict__ counters) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlock / 64);
  uint64_t found_total = 0;
  for (int64_t base = ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * span; base < words;
       base += nwaves * span) {
    const int64_t wl = base + lane;
    const bool mine = lane < span && wl < words;
    const uint64_t cl = mine ? vis[wl] : ~UINT64_C(0);  // lanes beyond the span or the end read as visited
    uint64_t open = __ballot(cl != ~UINT64_C(0));        // words with an unvisited vertex
    if (mine && cl == ~UINT64_C(0)) vis_next[wl] = cl;
    while (open) {
      // the next two open words (identical in every lane)
      const int j0 = __builtin_ctzll(open);
      open &= open - 1;
      const bool has1 = open != 0;
      const int j1 = has1 ? __builtin_ctzll(open) : j0;
      if (has1) open &= open - 1;
      const uint64_t c0 = __shfl(cl, j0, 64);
      const uint64_t c1 = has1 ? __shfl(cl, j1, 64) : ~UINT64_C(0);
      const int64_t w0 = base + j0, w1 = base + j1;
      const uint32_t v0 = (uint32_t)(w0 * 64 + lane), v1 = (uint32_t)(w1 * 64 + lane);
      const bool a0 = !((c0 >> lane) & 1), a1 = !((c1 >> lane) & 1);
      const uint32_t f0 = a0 ? first_nbr[v0] : kNoNbr;
      const uint32_t f1 = a1 ? first_nbr[v1] : kNoNbr;
      const bool h0 = f0 != kNoNbr && bit_set(vis, f0);
      const bool h1 = f1 != kNoNbr && bit_set(vis, f1);
      uint32_t p0 = f0, p1 = f1;
      // every lane makes both calls (the long-row scan needs the whole wave)
      const bool r0 = row_marked_nbr(f0 != kNoNbr && !h0, v0, vis, offsets, cols, p0);
      const bool r1 = row_marked_nbr(f1 != kNoNbr && !h1, v1, vis, offsets, cols, p1);
      const bool g0 = h0 || r0, g1 = h1 || r1;
      const uint64_t fm0 = __ballot(g0), fm1 = __ballot(g1);
      if (lane == 0) {
        vis_next[w0] = c0 | fm0;
        if (has1) vis_next[w1] = c1 | fm1;
      }
      if (g0) {
        level[v0] = (uint8_t)lvl;
        parent[v0] = p0;
      }
      if (g1) {
        level[v1] = (uint8_t)lvl;
        parent[v1] = p1;
      }
      found_total += __popcll(fm0) + __popcll(fm1);  // identical in every lane
    }
  }
  block_add(lane == 0 ? found_total : 0, counters + C_FOUND);
}

// Frontier bitmap -> frontier queue in ascending vertex order: an exclusive
// scan of per-word popcounts gives every word its output position, so no
// atomics are involved.  The bits are 'now & ~before' (before may be null).
struct FrontierBits {
  const uint64_t* now;
  const uint64_t* before;
  __device__ uint64_t operator()(size_t w) const {
    const uint64_t x = before ? (now[w] & ~before[w]) : now[w];
    return (uint64_t)__popcll(x);
  }
};
struct FrontierEmit {
  const uint64_t* now;
  const uint64_t* before;
  uint32_t* queue;
  __device__ void operator()(size_t w, uint64_t pre) const {
    uint64_t x = before ? (now[w] & ~before[w]) : now[w];
    while (x) {
      queue[pre++] = (uint32_t)(w * 64 + __builtin_ctzll(x));
      x &= x - 1;
    }
  }
};

// --- algorithmic byte model, evaluated from the final levels -------------
// top-down level L reads the rows of the vertices at level L-1
__global__ void __launch_bounds__(kBlock) k_model_td(const uint8_t* __restrict__ level,
                                                     const uint64_t* __restrict__ offsets, int64_t rows,
                                                     uint32_t L, unsigned long long* acc) {
  uint64_t b = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += (int64_t)gridDim.x * kBlock)
    if (level[v] == L - 1) b += 12 + 4 * (offsets[v + 1] - offsets[v]);
  block_add(b, acc);
}

// bottom-up level L probes, for every vertex with edges still unvisited, its
// neighbours in ascending order up to the first one visited before L
__global__ void __launch_bounds__(kBlock) k_model_bu(const uint8_t* __restrict__ level,
                                                     const uint64_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ cols, int64_t rows,
                                                     uint32_t L, unsigned long long* acc) {
  uint64_t b = 0;
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < rows; v += (int64_t)gridDim.x * kBlock) {
    const uint8_t lv = level[v];
    if (lv != 0xFF && lv < L) continue;
    const uint64_t s = offsets[v], e = offsets[v + 1];
    if (s == e) continue;
    uint64_t k = 0;
    for (uint64_t j = s; j < e; ++j) {
      ++k;
      const uint8_t lw = level[cols[j]];
      if (lw != 0xFF && lw < L) break;
    }
    b += 8 + 4 * k;
  }
  block_add(b, acc);
}

void bfs_model_bytes(g5_ctx* ctx, double* per_level, double* total) {
  Graph& g = ctx->graph;
  BfsState& b = ctx->bfs;
  G5_CHECK(b.have_result, "g5_bfs_model_bytes: no BFS result (call g5_run_bfs first)");
  const g5_bfs_stats& st = b.stats;
  hipStream_t s = ctx->stream;
  ctx->acc.alloc(4);
  const int64_t rows = g.n_rows;
  double tot = 1.0;  // the root's level byte
  for (int L = 0; L < G5_MAX_LEVELS; ++L)
    if (per_level) per_level[L] = 0;
  for (int L = 1; L < st.num_levels; ++L) {
    G5_HIP(hipMemsetAsync(ctx->acc.p, 0, 8, s));
    const unsigned grid = grid_of(rows, kBlock, 8192);
    if (st.direction[L])
      hipLaunchKernelGGL(k_model_td, dim3(grid), dim3(kBlock), 0, s, b.level.p, g.offsets.p, rows,
                         (uint32_t)L, (unsigned long long*)ctx->acc.p);
    else
      hipLaunchKernelGGL(k_model_bu, dim3(grid), dim3(kBlock), 0, s, b.level.p, g.offsets.p, g.cols.p,
                         rows, (uint32_t)L, (unsigned long long*)ctx->acc.p);
    G5_HIP(hipGetLastError());
    uint64_t h = 0;
    G5_HIP(hipMemcpyAsync(&h, ctx->acc.p, 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    const double nq = (double)st.nq[L];
    double bytes = (double)h + nq;  // + 1 B level write per discovered vertex
    if (st.direction[L]) bytes += 4.0 * nq;             // queue write
    else bytes += 3.0 * (double)g.n / 8.0;              // visited read, next read+write
    if (per_level) per_level[L] = bytes;
    tot += bytes;
  }
  if (total) *total = tot;
}

__global__ void k_make_pred(const uint32_t* __restrict__ old2new, const uint32_t* __restrict__ new2old,
                            const uint8_t* __restrict__ level, const uint32_t* __restrict__ parent,
                            int64_t n, int64_t rows, int64_t root_orig, int64_t* __restrict__ pred) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t nv = old2new[v];
    int64_t out = -1;
    if (v == root_orig) {
      out = root_orig;
    } else if ((int64_t)nv < rows) {
      const uint8_t l = level[nv];
      if (l != 0xFF) out = (int64_t)new2old[parent[nv]] | ((int64_t)l << 48);
    }
    pred[v] = out;
  }
}

__global__ void k_make_levels(const uint32_t* __restrict__ old2new, const uint8_t* __restrict__ level,
                              int64_t n, int64_t rows, int64_t root_orig, uint8_t* __restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t nv = old2new[v];
    uint8_t l = 0xFF;
    if (v == root_orig) l = 0;
    else if ((int64_t)nv < rows) l = level[nv];
    out[v] = l;
  }
}

// --- fused small top-down levels ---------------------------------------------
// Consecutive top-down levels whose frontiers are small (at most kFuseNf
// vertices with at most max_edges edges) run inside one launch of one
// workgroup: per level, the frontier's degree prefix in LDS, its edges taken
// kFuseThreads at a time (a lane finds its row by binary search in LDS), the
// visited bitmap test-and-set (plain-load filter, then atomicOr) and the next
// queue; between levels the workgroup applies the reference's switch rule
// (mpi/bfs.hpp:2799-2829: a growing top-down level that finds more vertices
// than the previous one and more than n_active/2000 hands over to bottom-up),
// so these levels cost no host round trip and no further launches.  The
// launch stops when a level finds nothing, before a level that pulls, before
// a frontier too large for it, or at max_levels.  Each level's vertex count
// and, at the end, the launch's state are published to coherent host memory,
// every value tagged with the launch's sequence number.
constexpr int kFuseThreads = 1024;
constexpr int kFuseNf = 4096;
constexpr int kFuseMaxLevels = 64;
enum { F_LEVELS = kFuseMaxLevels, F_STOP, F_QI, F_NF, F_E, F_NUM };
enum { FUSE_EMPTY = 0, FUSE_PULL = 1, FUSE_LARGE = 2, FUSE_LIMIT = 3 };

__global__ void __launch_bounds__(kFuseThreads) k_td_fused(
    uint32_t* __restrict__ q0, uint32_t* __restrict__ q1, int qi0, int64_t nf0,
    const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ cols, uint64_t* vis,
    uint8_t* __restrict__ level, uint32_t* __restrict__ parent, uint32_t lvl0, int max_levels, int64_t max_edges,
    double na, int64_t prev_nq0, int growing, Published* out, uint64_t seq) {
  __shared__ uint64_t s_pre[kFuseNf + 1];
  __shared__ uint32_t s_src[kFuseNf];
  __shared__ uint64_t s_wave[kFuseThreads / 64];
  __shared__ uint32_t s_cnt;
  constexpr int kPer = kFuseNf / kFuseThreads;  // frontier entries per thread in the scan
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int qi = qi0;
  int64_t nf = nf0, prev_nq = prev_nq0;
  uint64_t stop = FUSE_LIMIT, E = 0;
  int k = 0;
  while (k < max_levels) {
    const uint32_t* cur = qi ? q1 : q0;
    uint32_t* nxt = qi ? q0 : q1;
    if (nf > kFuseNf) {
      stop = FUSE_LARGE;
      E = 0;  // not known here
      break;
    }
    // degrees of the frontier, exclusive prefix in LDS (kPer per thread)
    uint64_t d[kPer], sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int64_t i = (int64_t)t * kPer + j;
      d[j] = 0;
      if (i < nf) {
        const uint32_t u = cur[i];
        s_src[i] = u;
        d[j] = offsets[u + 1] - offsets[u];
      }
      sum += d[j];
    }
    uint64_t inc = sum;  // inclusive scan of the per-thread sums within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t x = __shfl_up(inc, o, 64);
      if (lane >= o) inc += x;
    }
    if (lane == 63) s_wave[wave] = inc;
    if (t == 0) s_cnt = 0;
    __syncthreads();
    uint64_t base = 0;
    for (int w = 0; w < wave; ++w) base += s_wave[w];
    uint64_t run = base + inc - sum;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int64_t i = (int64_t)t * kPer + j;
      if (i < nf) s_pre[i] = run;
      run += d[j];
    }
    if (t == kFuseThreads - 1) s_pre[nf] = run;
    __syncthreads();
    E = s_pre[nf];
    if ((int64_t)E > max_edges) {
      stop = FUSE_LARGE;
      break;
    }
    const uint32_t lvl = lvl0 + (uint32_t)k;
    for (uint64_t e = t; e < E; e += kFuseThreads) {
      int lo = 0, hi = (int)nf - 1;  // last row starting at or before e
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= e) lo = mid; else hi = mid - 1;
      }
      const uint32_t u = s_src[lo];
      const uint32_t w = cols[offsets[u] + (e - s_pre[lo])];
      const uint64_t bit = UINT64_C(1) << (w & 63);
      if (vis[w >> 6] & bit) continue;
      const uint64_t old = atomicOr((unsigned long long*)(vis + (w >> 6)), (unsigned long long)bit);
      if (old & bit) continue;
      level[w] = (uint8_t)lvl;
      parent[w] = u;
      nxt[atomicAdd(&s_cnt, 1u)] = w;
    }
    __syncthreads();
    const int64_t nq = s_cnt;
    if (t == 0) publish(out + k, (uint64_t)nq, seq);
    const bool pull = growing && nq > prev_nq && (double)nq > na / 2000.0;
    prev_nq = nq;
    nf = nq;
    qi ^= 1;
    ++k;
    __syncthreads();  // every lane has read s_cnt and s_pre before the next level writes them
    if (nq == 0) {
      stop = FUSE_EMPTY;
      break;
    }
    if (pull) {
      stop = FUSE_PULL;
      break;
    }
  }
  if (t == 0) {
    publish(out + F_LEVELS, (uint64_t)k, seq);
    publish(out + F_STOP, stop, seq);
    publish(out + F_QI, (uint64_t)qi, seq);
    publish(out + F_NF, (uint64_t)nf, seq);
    publish(out + F_E, E, seq);
  }
}

// Level end: the counters the host decides the next level with (and the
// level's frontier edge count after a top-down level) are published straight
// into coherent host memory, each tagged with the level's sequence number, so
// the host need not wait for a copy and a stream synchronisation.
__global__ void k_publish(unsigned long long* __restrict__ counters, const uint64_t* __restrict__ prefix,
                          int64_t nf, int forward, Published* pub, uint64_t seq) {
  for (int i = 0; i < C_NUM; ++i) {
    const uint64_t v = (forward && i == 1) ? prefix[nf] : (uint64_t)counters[i];
    counters[i] = 0;  // ready for the next level, with no fill launched before it
    publish(pub + i, v, seq);
  }
}

// Waits for the level's counters (see PublishBlock::wait).
static void wait_published(BfsState& b, hipStream_t s) {
  b.pub.wait(0, C_NUM, b.pub_seq, s, "g5_run_bfs: level counters were not published");
  for (int i = 0; i < C_NUM; ++i) b.h_counters[i] = b.pub.value(i);
}

static void build_hub_index(g5_ctx* ctx, BfsState& b, hipStream_t s);

// Buffers of the segment-list levels, sized for the tuning so that no run
// allocates: a growing top-down level has at most n_active / 2000 frontier
// vertices (one more switches to bottom-up), so a head level's row suffixes
// take at most that many segments and a piece or block level that many
// times its pieces, within seg_max.  The block index of the first rows is
// built here too.  (The binned form, off by default, sizes its pair buffer
// by the level's edges inside the run.)
void bfs_tuning_buffers(g5_ctx* ctx) { bfs_tuning_buffers(ctx, ctx->bfs, ctx->stream); }

void bfs_tuning_buffers(g5_ctx* ctx, BfsState& b, hipStream_t s) {
  Graph& g = ctx->graph;
  if (b.hub_shift > 0 && b.hub_built_shift != b.hub_shift) build_hub_index(ctx, b, s);
  const int64_t seg_max = b.seg_max > 0 ? std::min<int64_t>(b.seg_max, kSegMax) : kSegMax;
  const int64_t nf_max = g.n_active / 2000 + 1;
  int64_t pieces = 1;
  if (b.td_pieces > 0) pieces = std::max<int64_t>(pieces, b.td_pieces);
  if (b.hub_shift > 0) pieces = std::max<int64_t>(pieces, b.hub_pieces);
  const size_t seg_n = (size_t)std::min<int64_t>(nf_max * pieces, seg_max) + 2;
  b.seg_pre.alloc(seg_n);
  b.seg_base.alloc(seg_n);
  b.seg_src.alloc(seg_n);
  b.seg_start.alloc(seg_n);
  b.seg_count.alloc(1);
  b.seg_scan_tmp.alloc(scan_tmp_elems(seg_n));
}

void bfs_prepare(g5_ctx* ctx) { bfs_prepare(ctx, ctx->bfs, ctx->stream); }

void bfs_prepare(g5_ctx* ctx, BfsState& b, hipStream_t s) {
  Graph& g = ctx->graph;
  G5_CHECK(g.built, "g5_prepare_bfs: no graph (call g5_construct first)");
  b.release();
  const int64_t rows = g.n_rows;
  b.words = (rows + 63) / 64;
  for (auto& v : b.vis) v.alloc(b.words);
  b.level.alloc(rows);
  b.parent.alloc(rows);
  for (auto& q : b.queue) q.alloc(rows);
  b.prefix.alloc(rows + 1);
  b.chunk_rows.alloc(2 * ((uint64_t)g.nnz / kEpb + 2));
  const size_t scan_n = (size_t)rows + 1;
  b.scan_tmp.alloc(scan_tmp_elems(scan_n));
  b.counters.alloc(C_NUM);
  bfs_tuning_buffers(ctx, b, s);
  G5_HIP(hipHostMalloc((void**)&b.h_counters, C_NUM * sizeof(uint64_t), hipHostMallocDefault));
  b.pub.alloc(C_NUM);
  b.pub_seq = 0;
  b.fuse.alloc(F_NUM);
  b.fuse_seq = 0;
  G5_HIP(hipEventCreate(&b.ev0));
  G5_HIP(hipEventCreate(&b.ev1));
  b.lev_begin.assign(G5_MAX_LEVELS, nullptr);
  b.lev_end.assign(G5_MAX_LEVELS, nullptr);
  for (int i = 0; i < G5_MAX_LEVELS; ++i) {
    G5_HIP(hipEventCreate(&b.lev_begin[i]));
    G5_HIP(hipEventCreate(&b.lev_end[i]));
  }
  b.prepared = true;
}

// The block index of the first rows: as many as fit kHubIndexBytes.
constexpr int64_t kHubIndexBytes = INT64_C(512) << 20;

static void build_hub_index(g5_ctx* ctx, BfsState& b, hipStream_t s) {
  Graph& g = ctx->graph;
  const int shift = (int)b.hub_shift;
  const int64_t nb = ((g.n_rows - 1) >> shift) + 1;
  b.hub_pieces = nb + 1;
  b.hub_rows = std::min<int64_t>(g.n_rows, kHubIndexBytes / (4 * b.hub_pieces));
  b.hub_bnd.alloc((size_t)(b.hub_rows * b.hub_pieces));
  hipLaunchKernelGGL(k_hub_bounds, dim3(grid_of(b.hub_rows * b.hub_pieces, kBlock, 65536)), dim3(kBlock), 0,
                     s, g.offsets.p, g.cols.p, b.hub_rows, b.hub_pieces, shift, b.hub_bnd.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipStreamSynchronize(s));
  b.hub_built_shift = b.hub_shift;
}

void bfs_run(g5_ctx* ctx, int64_t root) { bfs_run(ctx, ctx->bfs, ctx->stream, root); }

void bfs_run(g5_ctx* ctx, BfsState& b, hipStream_t s, int64_t root) {
  Graph& g = ctx->graph;
  G5_CHECK(b.prepared, "g5_run_bfs: call g5_prepare_bfs first");
  G5_CHECK(root >= 0 && root < g.n, "g5_run_bfs: root out of range");
  const int64_t rows = g.n_rows;
  unsigned long long* cnt = (unsigned long long*)b.counters.p;
  b.have_result = false;
  b.root_orig = root;
  g5_bfs_stats& st = b.stats;
  st = g5_bfs_stats{};

  if (b.hub_shift > 0 && b.hub_built_shift != b.hub_shift) bfs_tuning_buffers(ctx, b, s);

  G5_HIP(hipEventRecord(b.ev0, s));
  hipLaunchKernelGGL(k_bfs_init, dim3(grid_of(rows, kBlock, 4096)), dim3(kBlock), 0, s, b.level.p,
                     b.vis[0].p, rows, b.words);
  G5_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bfs_root, dim3(1), dim3(1), 0, s, g.old2new.p, root, rows, b.level.p,
                     b.parent.p, b.vis[0].p, b.queue[0].p, cnt);
  G5_HIP(hipGetLastError());

  // A root without edges (reachable only through a find_roots fallback) is
  // the whole tree: the reference's first top-down level finds nothing.
  const bool root_active = (g.has_edge_words[root >> 6] >> (root & 63)) & 1;
  int cur = 0, qi = 0;
  int64_t nf = 1;
  bool next_forward = true, growing = true;
  int64_t prev_nq = 0, visited = 1;
  // frontier edges of the coming top-down level when a small level summed
  // them (-1: unknown, both forms are launched and the device picks)
  int64_t e_next = -1;
  const double na = (double)g.n_active;
  int lvl = 0;
  if (!root_active) {
    st.num_levels = 2;
    st.direction[1] = 1;
    st.nq[1] = 0;
  }
  // a fused launch that found the next frontier too large hands that level
  // to the per-level path
  bool fuse_next = true;
  // A top-down level after the bottom-up phase whose frontier holds more
  // than 1/tail_pull as many vertices as remain unvisited (and more than a
  // fused launch takes) is found by the bottom-up kernel: an unvisited vertex
  // with a visited neighbour is on this level (an older neighbour would have
  // reached it before), so both kernels find the same vertices, and this one
  // probes from the few unvisited ones instead of expanding the frontier.
  // The level keeps its top-down place in the schedule.
  const auto pulls = [&](bool fwd, int64_t frontier) {
    return fwd && !growing && b.tail_pull > 0 && frontier > kFuseNf &&
           b.tail_pull * frontier > g.n_active - visited;
  };
  while (root_active) {
    ++lvl;
    G5_CHECK(lvl < 255, "g5_run_bfs: more than 254 levels (uint8 level array)");
    const bool forward = next_forward;
    G5_HIP(hipEventRecord(b.lev_begin[lvl], s));
    if (forward && fuse_next && b.small_edges > 0 && nf <= kFuseNf &&
        (e_next < 0 || e_next <= b.small_edges)) {
      // small top-down levels, as many as fit, in one single-workgroup launch
      const int max_levels = std::min(kFuseMaxLevels, 254 - lvl);
      ++b.fuse_seq;
      hipLaunchKernelGGL(k_td_fused, dim3(1), dim3(kFuseThreads), 0, s, b.queue[0].p, b.queue[1].p, qi, nf,
                         g.offsets.p, g.cols.p, b.vis[cur].p, b.level.p, b.parent.p, (uint32_t)lvl, max_levels,
                         b.small_edges, na, prev_nq, growing ? 1 : 0, b.fuse.d, b.fuse_seq);
      G5_HIP(hipGetLastError());
      G5_HIP(hipEventRecord(b.lev_end[lvl], s));
      const char* unpublished = "g5_run_bfs: fused levels were not published";
      b.fuse.wait(F_LEVELS, F_NUM, b.fuse_seq, s, unpublished);
      const int k = (int)b.fuse.value(F_LEVELS);
      b.fuse.wait(0, k, b.fuse_seq, s, unpublished);
      const uint64_t stop = b.fuse.value(F_STOP);
      if (k == 0) {
        // the first frontier was already too large: this level runs per level
        fuse_next = false;
        e_next = (int64_t)b.fuse.value(F_E);  // its edge count, when the launch summed it (0: not known)
        if (e_next == 0) e_next = -1;
        --lvl;
        continue;
      }
      for (int j = 0; j < k; ++j) {
        const int l = lvl + j;
        const int64_t nq = (int64_t)b.fuse.value(j);
        st.direction[l] = 1;
        st.nq[l] = nq;
        st.num_levels = l + 1;
        visited += nq;
        if (j > 0) {
          // the launch's time is the first fused level's; the others show 0
          G5_HIP(hipEventRecord(b.lev_begin[l], s));
          G5_HIP(hipEventRecord(b.lev_end[l], s));
        }
      }
      prev_nq = (int64_t)b.fuse.value(k - 1);
      qi = (int)b.fuse.value(F_QI);
      nf = (int64_t)b.fuse.value(F_NF);
      lvl += k - 1;
      if (stop == FUSE_EMPTY) break;
      next_forward = stop != FUSE_PULL;
      fuse_next = stop != FUSE_LARGE;
      e_next = stop == FUSE_LARGE && b.fuse.value(F_E) > 0 ? (int64_t)b.fuse.value(F_E) : -1;
      continue;
    }
    fuse_next = true;
    const bool pull_td = pulls(forward, nf);
    const bool td = forward && !pull_td;  // the top-down kernels run
    if (td) {
      device_scan(FrontierDegree{b.queue[qi].p, g.offsets.p, nf}, StoreU64{b.prefix.p},
                  (size_t)nf + 1, b.scan_tmp.p, s);
      if (b.bin_edges > 0) {
        // the binned form sizes its pair buffer from E: wait for it
        G5_HIP(hipMemcpyAsync(b.h_counters + 1, b.prefix.p + nf, 8, hipMemcpyDeviceToHost, s));
        G5_HIP(hipStreamSynchronize(s));
      }
      const uint64_t E_known = b.bin_edges > 0 ? b.h_counters[1] : 0;
      // past the bottom-up phase most targets are visited: a bulk level then
      // tests the visited bitmap before the level array
      const int probe_vis = (double)visited > na / 4 ? 1 : 0;
      // the level's frontier edge count when the previous level summed it
      const bool known_small = e_next >= 0 && (uint64_t)e_next <= kTdSmall;
      const bool known_bulk = e_next >= 0 && (uint64_t)e_next > kTdSmall;
      // while the search grows, a bulk level can take its rows block by
      // block (exact target blocks) or piece by piece (positional pieces)
      const int64_t seg_max = b.seg_max > 0 ? std::min<int64_t>(b.seg_max, kSegMax) : kSegMax;
      const bool blocked = b.hub_shift > 0 && !probe_vis && nf * b.hub_pieces <= seg_max;
      // a growing bulk level settles its highest-degree targets bottom-up
      // first (degree order only)
      bool head = b.td_head > 0 && !probe_vis && g.reorder_mode == 2;
      if (head && e_next > (int64_t)kTdSmall) {
        // E known on the host: no head when no row reaches the threshold
        // (the same test k_td_head_size makes)
        const double r = (double)g.nnz / (double)e_next;
        if ((double)b.td_head / 16.0 * r * r > (double)g.max_degree) head = false;
      }
      unsigned long long* head_words = head ? cnt + C_HEAD : nullptr;
      const bool pieces = blocked || (b.td_pieces > 0 && !probe_vis && nf * b.td_pieces <= seg_max);
      // with a head and row order, the expansion walks row suffixes (one
      // segment per frontier row: the segment count field bounds nf)
      if (head && !pieces && nf > seg_max) head = false;
      head_words = head ? cnt + C_HEAD : nullptr;
      const bool suffix = head && !pieces;
      const bool binned = b.bin_edges > 0 && E_known > kTdSmall && E_known >= (uint64_t)b.bin_edges;
      // Without binning no launch waits for E on the host: chunk rows and the
      // expansion take upper-bound grids and read E (prefix[nf]) on the
      // device; the small and the bulk expansion are both launched and the
      // one that does not fit E exits at once, as does k_td_finish after a
      // small level.  The host learns E with the level's counters.
      const uint64_t max_chunks = b.chunk_rows.n / 2;  // sized for every edge of the graph
      const uint64_t nchunks = b.bin_edges > 0 ? (E_known + kEpb - 1) / kEpb : max_chunks;
      const unsigned grid = b.bin_edges > 0 ? grid_of((int64_t)E_known, kEpb, 2048) : 2048u;
      if (nchunks)
        hipLaunchKernelGGL(k_chunk_rows<kEpb>,
                           dim3(b.bin_edges > 0 ? grid_of((int64_t)(2 * nchunks), kBlock, 65536) : 4096u),
                           dim3(kBlock), 0, s, b.prefix.p, nf, nchunks, b.chunk_rows.p);
      if (!known_bulk)
        hipLaunchKernelGGL(k_td_expand<false>, dim3(grid), dim3(kBlock), 0, s, b.queue[qi].p, nf, b.prefix.p,
                           b.chunk_rows.p, g.offsets.p, g.cols.p, b.vis[cur].p, b.level.p, b.parent.p,
                           b.queue[qi ^ 1].p, cnt, (uint32_t)lvl, 0, nullptr);
      if (head && !binned && !known_small) {
        hipLaunchKernelGGL(k_td_head_size, dim3(1), dim3(1), 0, s, b.prefix.p, nf, g.offsets.p, rows, b.words,
                           (double)g.nnz, (double)b.td_head / 16.0, head_words);
        hipLaunchKernelGGL(k_td_head, dim3(grid_of(b.words, kBlock / 64, 8192)), dim3(kBlock), 0, s,
                           b.vis[cur].p, g.first_nbr.p, g.offsets.p, g.cols.p, b.level.p, b.parent.p, head_words,
                           (uint32_t)lvl);
        G5_HIP(hipGetLastError());
      }
      if (known_small) {
        // nothing else to launch: the small expansion keeps the queue exact
      } else if (binned) {
        // binned by target range (see k_td_bin_count)
        const uint64_t E = E_known;
        int shift = kBinShift;
        while ((rows >> shift) >= kBinMaxBuckets) ++shift;
        const int nb = (int)(((rows - 1) >> shift) + 1);
        b.bin_pairs.alloc(E);
        b.bin_cnt.alloc(2 * kBinMaxBuckets);
        unsigned long long* bc = (unsigned long long*)b.bin_cnt.p;
        G5_HIP(hipMemsetAsync(bc, 0, nb * 8, s));
        hipLaunchKernelGGL(k_td_bin_count, dim3(grid), dim3(kBlock), 0, s, b.queue[qi].p, nf, b.prefix.p,
                           b.chunk_rows.p, g.offsets.p, g.cols.p, shift, nb, bc);
        hipLaunchKernelGGL(k_td_bin_offsets, dim3(1), dim3(kBlock), 0, s, bc, nb, bc + kBinMaxBuckets);
        hipLaunchKernelGGL(k_td_bin_scatter, dim3(grid), dim3(kBlock), 0, s, b.queue[qi].p, nf, b.prefix.p,
                           b.chunk_rows.p, g.offsets.p, g.cols.p, shift, nb, bc + kBinMaxBuckets, b.bin_pairs.p);
        hipLaunchKernelGGL(k_td_bin_apply, dim3(grid_of((int64_t)E, kBlock, 2048)), dim3(kBlock), 0, s,
                           b.bin_pairs.p, E, b.level.p, b.parent.p, (uint32_t)lvl);
      } else if (pieces || suffix) {
        // segment-list bulk level; a small level makes an empty segment list
        // and the expansion exits (both read E on the device)
        const size_t np = (size_t)(suffix ? nf : nf * (blocked ? b.hub_pieces : b.td_pieces)) + 1;
        b.seg_pre.alloc(np);
        b.seg_base.alloc(np);
        b.seg_src.alloc(np);
        b.seg_count.alloc(1);
        b.seg_scan_tmp.alloc(scan_tmp_elems(np));
        if (suffix) {
          b.seg_start.alloc(np);
          hipLaunchKernelGGL(k_suffix_start, dim3(grid_of(nf, kBlock, 4096)), dim3(kBlock), 0, s, b.queue[qi].p, nf,
                             b.prefix.p, g.offsets.p, g.cols.p, head_words, b.seg_start.p);
          const SuffixLen sl{b.queue[qi].p, g.offsets.p, b.prefix.p, nf, 1, b.seg_start.p};
          device_scan(sl, PieceEmit<SuffixLen>{sl, b.seg_pre.p, b.seg_base.p, b.seg_src.p, b.seg_count.p}, np,
                      b.seg_scan_tmp.p, s);
        } else if (blocked) {
          const BlockLen bl{b.queue[qi].p, g.offsets.p, b.prefix.p, nf, b.hub_pieces, b.hub_bnd.p, b.hub_rows};
          device_scan(bl, PieceEmit<BlockLen>{bl, b.seg_pre.p, b.seg_base.p, b.seg_src.p, b.seg_count.p}, np,
                      b.seg_scan_tmp.p, s);
        } else {
          const PieceLen pl{b.queue[qi].p, g.offsets.p, b.prefix.p, nf, b.td_pieces};
          device_scan(pl, PieceEmit<PieceLen>{pl, b.seg_pre.p, b.seg_base.p, b.seg_src.p, b.seg_count.p}, np,
                      b.seg_scan_tmp.p, s);
        }
        hipLaunchKernelGGL(k_chunk_segs, dim3(4096), dim3(kBlock), 0, s, b.seg_pre.p, b.seg_count.p, max_chunks,
                           b.chunk_rows.p);
        hipLaunchKernelGGL(k_td_expand_segs, dim3(grid), dim3(kBlock), 0, s, b.prefix.p, nf, b.seg_pre.p,
                           b.seg_base.p, b.seg_src.p, b.seg_count.p, b.chunk_rows.p, g.cols.p, b.vis[cur].p,
                           b.level.p, b.parent.p, (uint32_t)lvl, head_words);
      } else {
        hipLaunchKernelGGL(k_td_expand<true>, dim3(grid), dim3(kBlock), 0, s, b.queue[qi].p, nf, b.prefix.p,
                           b.chunk_rows.p, g.offsets.p, g.cols.p, b.vis[cur].p, b.level.p, b.parent.p,
                           (uint32_t*)nullptr, cnt, (uint32_t)lvl, probe_vis, binned ? nullptr : head_words);
      }
      G5_HIP(hipGetLastError());
      // after a bulk level, k_td_finish makes vis[cur^1] exactly this level's
      // vertices and adds them to vis[cur]
      if (!known_small)
        hipLaunchKernelGGL(k_td_finish, dim3(grid_of((b.words + 15) / 16, kBlock / 64, 8192)), dim3(kBlock), 0,
                           s, b.level.p, rows, b.words, (uint32_t)lvl, b.vis[cur].p, b.vis[cur ^ 1].p, cnt,
                           b.prefix.p, nf);
    } else {
      // words per wave step: as many as keep at least 32768 wave steps
      int span = 2;
      while (span < 64 && b.words / (2 * span) >= 32768) span *= 2;
      hipLaunchKernelGGL(k_bu, dim3(grid_of((b.words + span - 1) / span, kBlock / 64, 8192)), dim3(kBlock), 0, s,
                         b.vis[cur].p, b.vis[cur ^ 1].p, g.first_nbr.p, g.offsets.p, g.cols.p,
                         b.level.p, b.parent.p, b.words, span, (uint32_t)lvl, cnt);
    }
    G5_HIP(hipGetLastError());
    G5_HIP(hipEventRecord(b.lev_end[lvl], s));
    ++b.pub_seq;
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(1), 0, s, cnt, b.prefix.p, nf, td ? 1 : 0, b.pub.d,
                       b.pub_seq);
    G5_HIP(hipGetLastError());
    wait_published(b, s);
    const int64_t nq = (int64_t)b.h_counters[0];
    const bool bulk = td && b.h_counters[1] > kTdSmall;
    // a small top-down level summed its winners' degrees
    e_next = (td && !bulk) ? (int64_t)b.h_counters[C_NEXT_E] : -1;
    st.direction[lvl] = forward ? 1 : 0;
    st.bu_kernel[lvl] = td ? 0 : 1;
    st.nq[lvl] = nq;
    st.num_levels = lvl + 1;
    visited += nq;
    if (nq == 0) break;
    // direction switch, mpi/bfs.hpp:2799-2829
    const int64_t unvisited = g.n_active - visited;
    if (growing && nq > prev_nq) {
      if (forward && (double)nq > na / 2000.0) next_forward = false;
    } else {
      if (!forward && unvisited < (int64_t)(na / 8.0)) {
        next_forward = true;
        growing = false;
      }
    }
    prev_nq = nq;
    // The next top-down level needs this level's vertices as a queue, unless
    // the bottom-up kernel takes it.  A small top-down level appended them to
    // queue[qi^1]; a bulk one left them in vis[cur^1]; the bottom-up kernel
    // left them as vis[cur^1] & ~vis[cur].
    if (td && !bulk) {
      qi ^= 1;
    } else if (next_forward && !pulls(true, nq)) {
      const uint64_t* before = td ? nullptr : b.vis[cur].p;
      device_scan(FrontierBits{b.vis[cur ^ 1].p, before},
                  FrontierEmit{b.vis[cur ^ 1].p, before, b.queue[qi].p}, (size_t)b.words, b.scan_tmp.p, s);
    }
    if (!td) cur ^= 1;
    nf = nq;
  }
  G5_HIP(hipEventRecord(b.ev1, s));
  G5_HIP(hipEventSynchronize(b.ev1));
  float ms = 0;
  G5_HIP(hipEventElapsedTime(&ms, b.ev0, b.ev1));
  st.device_ms = ms;
  for (int l = 1; l <= lvl && root_active; ++l) {
    float lms = 0;
    G5_HIP(hipEventElapsedTime(&lms, b.lev_begin[l], b.lev_end[l]));
    st.level_ms[l] = lms;
  }
  st.reached = visited;
  b.have_result = true;
}

void bfs_get_pred(g5_ctx* ctx, int64_t* host_pred) {
  Graph& g = ctx->graph;
  BfsState& b = ctx->bfs;
  G5_CHECK(b.have_result, "g5_get_pred: no BFS result (call g5_run_bfs first)");
  b.pred_out.alloc(g.n);
  hipLaunchKernelGGL(k_make_pred, dim3(grid_of(g.n, kBlock, 65536)), dim3(kBlock), 0, ctx->stream,
                     g.old2new.p, g.new2old.p, b.level.p, b.parent.p, g.n, g.n_rows, b.root_orig,
                     b.pred_out.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipMemcpyAsync(host_pred, b.pred_out.p, g.n * 8, hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(hipStreamSynchronize(ctx->stream));
}

void bfs_get_levels(g5_ctx* ctx, uint8_t* host_levels) {
  Graph& g = ctx->graph;
  BfsState& b = ctx->bfs;
  G5_CHECK(b.have_result, "g5_get_levels: no BFS result (call g5_run_bfs first)");
  b.level_out.alloc(g.n);
  hipLaunchKernelGGL(k_make_levels, dim3(grid_of(g.n, kBlock, 65536)), dim3(kBlock), 0, ctx->stream,
                     g.old2new.p, b.level.p, g.n, g.n_rows, b.root_orig, b.level_out.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipMemcpyAsync(host_levels, b.level_out.p, g.n, hipMemcpyDeviceToHost, ctx->stream));
  G5_HIP(hipStreamSynchronize(ctx->stream));
}

int64_t bfs_teps_edges(g5_ctx* ctx) {
  ctx->acc.alloc(4);
  return bfs_teps_edges(ctx, ctx->bfs, ctx->acc, ctx->stream);
}

int64_t bfs_teps_edges(g5_ctx* ctx, const BfsState& b, DevBuf<uint64_t>& acc, hipStream_t s) {
  Graph& g = ctx->graph;
  G5_CHECK(b.have_result, "g5_teps_edges: no BFS result (call g5_run_bfs first)");
  G5_CHECK(acc.n >= 4, "g5_teps_edges: reduction buffer too small");
  const uint64_t sum = device_reduce_sum(ReachedWeight{b.level.p, g.weight.p}, (size_t)g.n_rows, acc.p, s);
  return (int64_t)(sum / 2);
}

}  // namespace g5
