// g5_ms_device.hpp -- device steps of the batched (64 roots per word) BFS,
// shared by the batch kernels (g5_msbfs.hip) and the serving engine
// (g5_serve.hip).  Every step takes its workgroup index and count, so the
// same code runs as a grid-stride kernel or inside the persistent kernel.
#pragma once

#include "g5_frontier.hpp"

namespace g5 {

constexpr uint32_t kMsNoNbr = 0xFFFFFFFFu;
constexpr int kMsShortRow = 32;

enum { M_LIVE = 0, M_PAIRS = 1, M_NZ = 2, M_FDEG = 3, M_NUM = 8 };

__device__ __forceinline__ uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Block-wide OR / sums of the level counters, one atomic each per workgroup.
__device__ __forceinline__ void ms_block_counters(uint64_t live, uint64_t pairs, uint64_t nz, uint64_t fdeg,
                                                  unsigned long long* c) {
  __shared__ uint64_t s_v[4][kBlock / 64];
  live = wave_or(live);
  pairs = wave_sum(pairs);
  nz = wave_sum(nz);
  fdeg = wave_sum(fdeg);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_v[0][wave] = live;
    s_v[1][wave] = pairs;
    s_v[2][wave] = nz;
    s_v[3][wave] = fdeg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, b = 0, d = 0, e = 0;
    for (int i = 0; i < kBlock / 64; ++i) {
      a |= s_v[0][i];
      b += s_v[1][i];
      d += s_v[2][i];
      e += s_v[3][i];
    }
    if (a) atomicOr(c + M_LIVE, (unsigned long long)a);
    if (b) atomicAdd(c + M_PAIRS, (unsigned long long)b);
    if (d) atomicAdd(c + M_NZ, (unsigned long long)d);
    if (e) atomicAdd(c + M_FDEG, (unsigned long long)e);
  }
  __syncthreads();
}


// Vertices of the frontier union (set bits of the nonzero bitmap) -> queue,
// ascending, through a scan of per-word popcounts.
struct UnionCount {
  const uint64_t* nz;
  __device__ uint64_t operator()(size_t w) const { return (uint64_t)__popcll(nz[w]); }
};
// The same over words [0, words) with one more element whose prefix, the
// total, goes to *count.
struct BitsCountN {
  const uint64_t* bits;
  int64_t words;
  __device__ uint64_t operator()(size_t w) const { return (int64_t)w < words ? (uint64_t)__popcll(bits[w]) : 0; }
};
struct BitsEmitN {
  const uint64_t* bits;
  int64_t words;
  uint32_t* list;
  uint64_t* count;
  __device__ void operator()(size_t w, uint64_t pre) const {
    if ((int64_t)w == words) {
      *count = pre;
      return;
    }
    uint64_t x = bits[w];
    while (x) {
      list[pre++] = (uint32_t)(w * 64 + __builtin_ctzll(x));
      x &= x - 1;
    }
  }
};
struct UnionEmit {
  const uint64_t* nz;
  uint32_t* queue;
  __device__ void operator()(size_t w, uint64_t pre) const {
    uint64_t x = nz[w];
    while (x) {
      queue[pre++] = (uint32_t)(w * 64 + __builtin_ctzll(x));
      x &= x - 1;
    }
  }
};


// Per-level hooks of the finish and pull steps.  keep: lanes of seen that
// stay (the others belong to finished searches and are cleared on the way);
// epi(v, inb, nw): called by every lane of a wave with its vertex's new lanes;
// epi.begin() / epi.end() by every thread of the block around the sweep.
struct NoEpilogue {
  __device__ void begin() const {}
  __device__ void operator()(int64_t, bool, uint64_t) const {}
  __device__ void end() const {}
};

// Per-lane statistics of a level: for every lane the new vertices and the sum
// of their row lengths (the lane's next frontier edges).  Thread i of a wave
// keeps lane i's sums in registers: per 64-vertex word it walks the vertices
// with new lanes, reading each one's word and row length by readlane, so no
// two threads add to one place until the block's LDS reduction at the end;
// totals go to lane_nq[0, 64) and lane_nq[64, 128).
struct LaneEpi {
  unsigned long long* lane_nq;
  const uint64_t* offsets;
  uint64_t stat;  // the lanes whose statistics are kept (the ones still choosing a direction)
  mutable uint64_t n_acc = 0, d_acc = 0;
  __device__ unsigned long long* lds() const {
    __shared__ unsigned long long s_lane[128];
    return s_lane;
  }
  __device__ void begin() const {
    if (threadIdx.x < 128) lds()[threadIdx.x] = 0;
    __syncthreads();
  }
  __device__ void operator()(int64_t v, bool inb, uint64_t nw_all) const {
    const uint64_t nw = nw_all & stat;
    const uint64_t deg = (inb && nw) ? offsets[v + 1] - offsets[v] : 0;
    const int me = threadIdx.x & 63;
    uint64_t m = __ballot(inb && nw != 0);
    while (m) {
      const int t = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const uint32_t wlo = __builtin_amdgcn_readlane((uint32_t)nw, t);
      const uint32_t whi = __builtin_amdgcn_readlane((uint32_t)(nw >> 32), t);
      const uint32_t dlo = __builtin_amdgcn_readlane((uint32_t)deg, t);
      const uint32_t dhi = __builtin_amdgcn_readlane((uint32_t)(deg >> 32), t);
      const uint32_t bit = me < 32 ? (wlo >> me) & 1 : (whi >> (me - 32)) & 1;
      if (bit) {
        ++n_acc;
        d_acc += ((uint64_t)dhi << 32) | dlo;
      }
    }
  }
  __device__ void end() const {
    const int me = threadIdx.x & 63;
    if (n_acc) {
      atomicAdd(&lds()[me], (unsigned long long)n_acc);
      atomicAdd(&lds()[64 + me], (unsigned long long)d_acc);
    }
    __syncthreads();
    if (threadIdx.x < 128 && lds()[threadIdx.x]) atomicAdd(lane_nq + threadIdx.x, lds()[threadIdx.x]);
  }
};

// After a push or a sweep: next keeps only lanes new to v, seen gains them,
// and nz_next gets one bit per vertex whose frontier word is not zero.
// Blocks bid of nblk take 64-vertex words in a grid-stride loop.
template <class Epi = NoEpilogue>
__device__ __forceinline__ void ms_finish_body(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                               uint64_t* __restrict__ nz_next, const uint64_t* __restrict__ offsets,
                                               int64_t rows, unsigned long long* __restrict__ c, unsigned bid,
                                               unsigned nblk, uint64_t keep = ~UINT64_C(0), const Epi& epi = Epi()) {
  uint64_t live = 0, pairs = 0, nz = 0, fdeg = 0;
  const int64_t words = (rows + 63) / 64;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)nblk * (kBlock / 64);
  epi.begin();
  for (int64_t wd = (int64_t)bid * (kBlock / 64) + (threadIdx.x >> 6); wd < words; wd += nwaves) {
    const int64_t v = wd * 64 + lane;
    const bool inb = v < rows;
    uint64_t nw = 0;
    if (inb) {
      const uint64_t s0 = seen[v];
      const uint64_t s = s0 & keep;
      const uint64_t x = next[v];
      nw = x & ~s;
      if (nw != x) next[v] = nw;  // most words are already right (zero or all new)
      if (nw || s != s0) seen[v] = s | nw;
      if (nw) {
        live |= nw;
        pairs += __popcll(nw);
        ++nz;
        fdeg += offsets[v + 1] - offsets[v];
      }
    }
    epi(v, inb, nw);
    const uint64_t b = __ballot(nw != 0);
    if (lane == 0) nz_next[wd] = b;
  }
  epi.end();
  ms_block_counters(live, pairs, nz, fdeg, c);
}


// The same fold over a list of the vertices a sparse push touched (each once,
// in any order): 64 list entries per wave and step; nz_next (cleared
// beforehand) gets its bits by atomicOr since a word's vertices may sit in
// different waves.
template <class Epi = NoEpilogue>
__device__ __forceinline__ void ms_finish_list_body(uint64_t* __restrict__ seen, uint64_t* __restrict__ next,
                                                    uint64_t* __restrict__ nz_next,
                                                    const uint64_t* __restrict__ offsets,
                                                    const uint32_t* __restrict__ list, uint64_t n,
                                                    unsigned long long* __restrict__ c, unsigned bid, unsigned nblk,
                                                    const Epi& epi = Epi()) {
  uint64_t live = 0, pairs = 0, nz = 0, fdeg = 0;
  const int lane = threadIdx.x & 63;
  const uint64_t nwaves = (uint64_t)nblk * (kBlock / 64);
  epi.begin();
  for (uint64_t base = ((uint64_t)bid * (kBlock / 64) + (threadIdx.x >> 6)) * 64; base < n; base += nwaves * 64) {
    const uint64_t i = base + lane;
    const bool inb = i < n;
    const int64_t v = inb ? (int64_t)list[i] : 0;
    uint64_t nw = 0;
    if (inb) {
      const uint64_t s0 = seen[v];
      const uint64_t x = next[v];
      nw = x & ~s0;
      if (nw != x) next[v] = nw;
      if (nw) {
        seen[v] = s0 | nw;
        live |= nw;
        pairs += __popcll(nw);
        ++nz;
        fdeg += offsets[v + 1] - offsets[v];
        atomicOr((unsigned long long*)(nz_next + (v >> 6)), (unsigned long long)(UINT64_C(1) << (v & 63)));
      }
    }
    epi(v, inb, nw);
  }
  epi.end();
  ms_block_counters(live, pairs, nz, fdeg, c);
}

// Pull step: one lane per vertex, the wave helps with long rows.
// How the pull reads a neighbour's frontier word:
//   kFrontPlain    front[w] (8-byte random read into an n_rows-word array)
//   kFrontFilter   only when w's bit in the nonzero bitmap is set (one more
//                  dependent load, far fewer 8-byte reads)
//   kFrontCompact  from the compact frontier: pack[2 * (w / 64)] holds the
//                  nonzero bits of w's 64-vertex word and pack[2 * (w / 64) + 1]
//                  the number of nonzero words before it, so the word is
//                  compact[rank + popcount(bits below w)]; the compact array
//                  holds only the frontier's words, small enough at the
//                  sparse levels to stay in the last-level cache
//   kFrontHybrid   plain for w < filter_from (the high-degree vertices, whose
//                  words are often in the frontier and in cache), filtered above
enum { kFrontPlain = 0, kFrontFilter = 1, kFrontCompact = 2, kFrontHybrid = 3 };

struct FrontView {
  const uint64_t* front;
  const uint64_t* nz;
  const uint64_t* pack;
  const uint64_t* compact;
  int mode;
  uint32_t filter_from;
};

__device__ __forceinline__ uint64_t front_of(const FrontView& f, uint32_t w) {
  if (f.mode == kFrontCompact) {
    const uint64_t bits = f.pack[2 * (w >> 6)];
    if (!((bits >> (w & 63)) & 1)) return 0;
    const uint64_t below = bits & ((UINT64_C(1) << (w & 63)) - 1);
    return f.compact[f.pack[2 * (w >> 6) + 1] + (uint64_t)__popcll(below)];
  }
  if (f.mode == kFrontFilter && !bit_set(f.nz, w)) return 0;
  if (f.mode == kFrontHybrid && w >= f.filter_from && !bit_set(f.nz, w)) return 0;
  return f.front[w];
}

// Compact frontier of a level (scan over the nonzero bitmap's words).
struct CompactCount {
  const uint64_t* nz;
  __device__ uint64_t operator()(size_t w) const { return (uint64_t)__popcll(nz[w]); }
};
struct CompactEmit {
  const uint64_t* nz;
  const uint64_t* front;
  uint64_t* pack;
  uint64_t* compact;
  __device__ void operator()(size_t w, uint64_t pre) const {
    uint64_t x = nz[w];
    pack[2 * w] = x;
    pack[2 * w + 1] = pre;
    while (x) {
      compact[pre++] = front[w * 64 + __builtin_ctzll(x)];
      x &= x - 1;
    }
  }
};

// The probes of one pull step for each lane's vertex v (need: the lanes v
// still has to find; 0 for a lane with nothing to do): its first neighbour,
// then its short row lane-sequentially or its long row by the whole wave.
// Every lane of the wave calls it.  Returns the OR of the frontier words
// probed (at least the needed lanes when they are all present).
template <int kU, int kW>
__device__ __forceinline__ uint64_t ms_pull_probe(const FrontView& fv, int64_t v, uint64_t need,
                                                  const uint32_t* __restrict__ first_nbr,
                                                  const uint64_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ cols) {
  const int lane = threadIdx.x & 63;
  uint64_t acc = 0;
  bool more = false;
  uint64_t rb = 0, re = 0;
  if (need) {
    const uint32_t f = first_nbr[v];
    if (f != kMsNoNbr) {
      acc = front_of(fv, f);
      if ((acc & need) != need) {
        rb = offsets[v] + 1;
        re = offsets[v + 1];
        more = rb < re;
      }
    }
  }
  if (more && re - rb <= (uint64_t)kMsShortRow) {
    // kU independent probes per step (entries past the row repeat the first)
    for (uint64_t k = rb; k < re; k += kU) {
      uint32_t w[kU];
      w[0] = cols[k];
#pragma unroll
      for (int j = 1; j < kU; ++j) w[j] = k + j < re ? cols[k + j] : w[0];
      uint64_t x = 0;
#pragma unroll
      for (int j = 0; j < kU; ++j) x |= front_of(fv, w[j]);
      acc |= x;
      if ((acc & need) == need) break;
    }
    more = false;
  }
  uint64_t lm = __ballot(more);
  while (lm) {
    const int src = __ffsll((unsigned long long)lm) - 1;
    lm &= lm - 1;
    const uint64_t b = __shfl(rb, src, 64), e = __shfl(re, src, 64);
    const uint64_t nd = __shfl(need, src, 64);
    uint64_t a = __shfl(acc, src, 64);
    // 64 kW entries per step, kW independent probes per lane
    for (uint64_t base = b; base < e; base += 64 * kW) {
      uint32_t w[kW];
#pragma unroll
      for (int j = 0; j < kW; ++j) w[j] = base + lane + 64 * j < e ? cols[base + lane + 64 * j] : 0;
      uint64_t x = 0;
#pragma unroll
      for (int j = 0; j < kW; ++j)
        if (base + lane + 64 * j < e) x |= front_of(fv, w[j]);
      a |= wave_or(x);
      if ((a & nd) == nd) break;
    }
    if (lane == src) acc = a;
  }
  return acc;
}

// kU: independent probes per step of a lane's short row; kW: entries per lane
// and step of the wave's scan of a long row (kW x 64 per step).
template <class Epi = NoEpilogue, int kU = 4, int kW = 2>
__device__ __forceinline__ void ms_pull_body(uint64_t* __restrict__ seen, const FrontView fv,
                                             uint64_t* __restrict__ next, uint64_t* __restrict__ nz_next,
                                             const uint32_t* __restrict__ first_nbr,
                                             const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ cols,
                                             int64_t rows, uint64_t live_in, unsigned long long* __restrict__ c,
                                             unsigned bid, unsigned nblk, uint64_t keep = ~UINT64_C(0),
                                             const Epi& epi = Epi(), bool merge = false,
                                             uint64_t* __restrict__ miss = nullptr) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)nblk * (kBlock / 64);
  uint64_t live = 0, pairs = 0, nz = 0, fdeg = 0;
  const int64_t words = (rows + 63) / 64;
  epi.begin();
  for (int64_t wd = (int64_t)bid * (kBlock / 64) + (threadIdx.x >> 6); wd < words; wd += nwaves) {
    const int64_t v = wd * 64 + lane;
    const bool inb = v < rows;
    const uint64_t s0 = inb ? seen[v] : ~UINT64_C(0);
    const uint64_t s = inb ? (s0 & keep) : s0;
    const uint64_t need = ~s & live_in;
    // lanes pushed into next earlier in this level (merge): new ones join nw
    const uint64_t old = (merge && inb) ? next[v] : 0;
    const uint64_t pushed = old & ~s;
    const uint64_t acc = ms_pull_probe<kU, kW>(fv, v, need, first_nbr, offsets, cols);
    const uint64_t nw = (acc & need) | pushed;
    epi(v, inb, nw);
    // one bit per vertex still missing a live lane, for the list-driven tail levels
    if (miss) {
      const uint64_t mb = __ballot(inb && (live_in & ~(s | nw)) != 0);
      if (lane == 0) miss[wd] = mb;
    }
    const uint64_t b = __ballot(nw != 0);
    if (lane == 0) nz_next[wd] = b;
    if (inb) {
      // with merge the array was cleared before the push and holds zeros or
      // pushed lanes: only words that change are written
      if (!merge || nw != old) next[v] = nw;
      if (nw || s != s0) seen[v] = s | nw;
      if (nw) {
        live |= nw;
        pairs += __popcll(nw);
        ++nz;
        fdeg += offsets[v + 1] - offsets[v];
      }
    }
  }
  epi.end();
  ms_block_counters(live, pairs, nz, fdeg, c);
}


// A pull step over a list of the vertices that may still miss a live lane
// (the tail levels of a batch): the level's frontier array, nz_next and miss
// were cleared beforehand, so only listed vertices with new lanes are
// written, and the ones still missing a live lane set their bit in miss.
template <int kU = 4, int kW = 2>
__device__ __forceinline__ void ms_pull_list_body(uint64_t* __restrict__ seen, const FrontView fv,
                                                  uint64_t* __restrict__ next, uint64_t* __restrict__ nz_next,
                                                  const uint32_t* __restrict__ first_nbr,
                                                  const uint64_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ cols, const uint32_t* __restrict__ list,
                                                  uint64_t n, uint64_t live_in, uint64_t* __restrict__ miss,
                                                  unsigned long long* __restrict__ c, unsigned bid, unsigned nblk) {
  const int lane = threadIdx.x & 63;
  const uint64_t nwaves = (uint64_t)nblk * (kBlock / 64);
  uint64_t live = 0, pairs = 0, nz = 0, fdeg = 0;
  for (uint64_t base = ((uint64_t)bid * (kBlock / 64) + (threadIdx.x >> 6)) * 64; base < n; base += nwaves * 64) {
    const uint64_t i = base + lane;
    const bool inb = i < n;
    const int64_t v = inb ? (int64_t)list[i] : 0;
    const uint64_t s = inb ? seen[v] : ~UINT64_C(0);
    const uint64_t need = ~s & live_in;
    const uint64_t acc = ms_pull_probe<kU, kW>(fv, v, need, first_nbr, offsets, cols);
    const uint64_t nw = acc & need;
    if (nw) {
      next[v] = nw;
      seen[v] = s | nw;
      live |= nw;
      pairs += __popcll(nw);
      ++nz;
      fdeg += offsets[v + 1] - offsets[v];
      atomicOr((unsigned long long*)(nz_next + (v >> 6)), (unsigned long long)(UINT64_C(1) << (v & 63)));
    }
    if (inb && (live_in & ~(s | nw)) != 0)
      atomicOr((unsigned long long*)(miss + (v >> 6)), (unsigned long long)(UINT64_C(1) << (v & 63)));
  }
  ms_block_counters(live, pairs, nz, fdeg, c);
}


// Dense pull step as an edge-parallel sweep over the whole CSR: each lane
// takes kSweepEpt consecutive column entries (a wave reads 2 KiB of cols in
// one coalesced pass), ORs the frontier words of the neighbours whose
// nonzero bit is set into its row's LDS accumulator, and the block writes
// one word per row: a plain store for rows that lie inside the chunk, an
// atomicOr for the (at most two) rows that cross its ends, whose words were
// cleared beforehand by k_ms_sweep_clear.  No early exit: this is the form
// for levels at which most rows would be scanned to the end anyway.  Rows
// are those of modes 1 and 2 (every row has at least one entry), so at most
// kSweepEpb rows overlap a chunk.  next then holds, per vertex, the lanes of
// its neighbours' frontier words; k_ms_finish keeps the new ones.
constexpr int kSweepEpt = 4;
constexpr int kSweepEpb = kSweepEpt * kBlock;

__device__ __forceinline__ void ms_sweep_clear_body(const uint32_t* __restrict__ chunk_rows, uint64_t nchunks,
                                                    const uint64_t* __restrict__ offsets, uint64_t* __restrict__ next,
                                                    unsigned bid, unsigned nblk) {
  for (uint64_t c = (uint64_t)bid * kBlock + threadIdx.x; c < nchunks; c += (uint64_t)nblk * kBlock) {
    const uint32_t r0 = chunk_rows[2 * c], r1 = chunk_rows[2 * c + 1];
    const uint64_t e0 = c * kSweepEpb;
    if (offsets[r0] < e0) next[r0] = 0;                  // first row began in an earlier chunk
    if (offsets[r1 + 1] > e0 + kSweepEpb) next[r1] = 0;  // last row goes on into a later chunk
  }
}

// LDS of one sweep chunk: row starts relative to the chunk's first entry e0
// (clamped to -1 before it and kSweepEpb + 1 after it, which keeps every
// comparison below exact) and row accumulators.
struct SweepLds {
  int32_t off[kSweepEpb + 2];
  unsigned long long acc[kSweepEpb + 1];
};

// Stages the row starts of chunk c; returns the row count (every lane calls
// it; ends with a barrier).
__device__ __forceinline__ int sweep_stage(uint64_t chunk, const uint32_t* __restrict__ chunk_rows,
                                           const uint64_t* __restrict__ offsets, SweepLds& L, bool clear_acc) {
  const uint32_t r0 = chunk_rows[2 * chunk];
  const int64_t e0 = (int64_t)(chunk * kSweepEpb);
  int nr = (int)(chunk_rows[2 * chunk + 1] - r0 + 1);
  nr = nr > kSweepEpb ? kSweepEpb : nr;  // bound for LDS; rows are non-empty in modes 1 and 2
  for (int i = threadIdx.x; i <= nr; i += kBlock) {
    int64_t d = (int64_t)offsets[r0 + i] - e0;
    d = d < -1 ? -1 : (d > kSweepEpb + 1 ? kSweepEpb + 1 : d);
    L.off[i] = (int32_t)d;
    if (clear_acc && i < nr) L.acc[i] = 0;
  }
  __syncthreads();
  return nr;
}

// Last staged row starting at or before relative entry e (LDS binary search).
__device__ __forceinline__ int sweep_row_of(const SweepLds& L, int nr, int e) {
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ void ms_sweep_body(const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                              const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ cols,
                                              const uint64_t* __restrict__ front,
                                              const uint64_t* __restrict__ nz_front, uint64_t* __restrict__ next,
                                              SweepLds& L, unsigned bid, unsigned nblk) {
  for (uint64_t chunk = bid; chunk * kSweepEpb < nnz; chunk += nblk) {
    const uint64_t e0 = chunk * kSweepEpb;
    const uint64_t e1 = (e0 + kSweepEpb < nnz) ? e0 + kSweepEpb : nnz;
    const uint32_t r0 = chunk_rows[2 * chunk];
    const int nr = sweep_stage(chunk, chunk_rows, offsets, L, true);
    const int len = (int)(e1 - e0);
    const int my0 = threadIdx.x * kSweepEpt;  // relative to e0
    if (my0 < len) {
      int lo = sweep_row_of(L, nr, my0);
      const int my1 = (my0 + kSweepEpt < len) ? my0 + kSweepEpt : len;
      uint32_t w[kSweepEpt];
#pragma unroll
      for (int j = 0; j < kSweepEpt; ++j) w[j] = (my0 + j < my1) ? cols[e0 + my0 + j] : 0;
      uint64_t x[kSweepEpt];
#pragma unroll
      for (int j = 0; j < kSweepEpt; ++j) x[j] = (my0 + j < my1 && bit_set(nz_front, w[j])) ? front[w[j]] : 0;
      uint64_t acc = 0;
      for (int j = 0; j < kSweepEpt && my0 + j < my1; ++j) {
        while (lo + 1 < nr && L.off[lo + 1] <= my0 + j) {
          if (acc) atomicOr(&L.acc[lo], (unsigned long long)acc);
          acc = 0;
          ++lo;
        }
        acc |= x[j];
      }
      if (acc) atomicOr(&L.acc[lo], (unsigned long long)acc);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nr; i += kBlock) {
      const uint64_t a = L.acc[i];
      if (L.off[i] >= 0 && L.off[i + 1] <= len)
        next[r0 + i] = a;
      else if (a)
        atomicOr((unsigned long long*)(next + r0 + i), (unsigned long long)a);
    }
    __syncthreads();
  }
}


// Push step over the CSR in sweep chunks (no frontier queue, so no scan): the
// chunks holding a frontier row are collected first from the nonzero bitmap;
// in those, consecutive lanes take consecutive entries and, for those whose
// row u is in the frontier, send front[u] to the target w.  Lanes that w has
// already seen are not filtered here (that would cost one more random read
// per entry): folding next into seen drops them.
__device__ __forceinline__ bool chunk_active(uint64_t chunk, const uint32_t* __restrict__ chunk_rows,
                                             const uint64_t* __restrict__ nz_front) {
  const uint32_t r0 = chunk_rows[2 * chunk], r1 = chunk_rows[2 * chunk + 1];
  for (uint32_t wd = r0 >> 6; wd <= (r1 >> 6); ++wd) {
    uint64_t x = nz_front[wd];
    if (wd == (r0 >> 6)) x &= ~UINT64_C(0) << (r0 & 63);
    if (wd == (r1 >> 6) && (r1 & 63) != 63) x &= (UINT64_C(1) << ((r1 & 63) + 1)) - 1;
    if (x) return true;
  }
  return false;
}

// One chunk of the push (every lane of the block calls it; LDS barriers).
__device__ __forceinline__ void push_chunk(uint64_t chunk, const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                           const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ cols,
                                           const uint64_t* __restrict__ front,
                                           const uint64_t* __restrict__ nz_front, uint64_t* next, SweepLds& L,
                                           uint64_t lanes = ~UINT64_C(0)) {
  const uint32_t r0 = chunk_rows[2 * chunk];
  const uint64_t e0 = chunk * kSweepEpb;
  const uint64_t e1 = (e0 + kSweepEpb < nnz) ? e0 + kSweepEpb : nnz;
  const int nr = sweep_stage(chunk, chunk_rows, offsets, L, false);
  const int len = (int)(e1 - e0);
  // consecutive lanes take consecutive entries (coalesced columns; targets
  // of a sorted long row close together); a lane re-finds its row only when
  // an entry lies past it
  int lo = -1;
  uint64_t fu = 0;
  for (int e = threadIdx.x; e < len; e += kBlock) {
    if (lo < 0 || (lo + 1 < nr && L.off[lo + 1] <= e)) {
      int a = lo < 0 ? 0 : lo + 1, b = nr - 1;  // last staged row starting at or before e
      while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (L.off[mid] <= e) a = mid; else b = mid - 1;
      }
      lo = a;
      const uint32_t u = r0 + lo;
      fu = bit_set(nz_front, u) ? front[u] & lanes : 0;
    }
    if (!fu) continue;
    const uint32_t w = cols[e0 + e];
    if (fu & ~__atomic_load_n(next + w, __ATOMIC_RELAXED))
      atomicOr((unsigned long long*)(next + w), (unsigned long long)fu);
  }
  __syncthreads();
}

// push_chunk that also lists the targets it touches: the one atomicOr that
// finds a target's word still zero (next was cleared for the level) appends
// the target to touch (wave-aggregated appends to *touch_count); every
// target is listed once, so touch needs room for the rows.
__device__ __forceinline__ void push_chunk_touch(uint64_t chunk, const uint32_t* __restrict__ chunk_rows,
                                                 uint64_t nnz, const uint64_t* __restrict__ offsets,
                                                 const uint32_t* __restrict__ cols, const uint64_t* __restrict__ front,
                                                 const uint64_t* __restrict__ nz_front, uint64_t* next, SweepLds& L,
                                                 uint64_t lanes, uint32_t* __restrict__ touch,
                                                 unsigned long long* __restrict__ touch_count) {
  const int lane = threadIdx.x & 63;
  const uint32_t r0 = chunk_rows[2 * chunk];
  const uint64_t e0 = chunk * kSweepEpb;
  const uint64_t e1 = (e0 + kSweepEpb < nnz) ? e0 + kSweepEpb : nnz;
  const int nr = sweep_stage(chunk, chunk_rows, offsets, L, false);
  const int len = (int)(e1 - e0);
  int lo = -1;
  uint64_t fu = 0;
  // every lane of a wave runs the same number of steps (kBlock is a multiple
  // of the wave), so the appends' ballots see the whole wave
  for (int e0l = threadIdx.x & ~63; e0l < len; e0l += kBlock) {
    const int e = e0l + lane;
    bool first = false;
    uint32_t w = 0;
    if (e < len) {
      if (lo < 0 || (lo + 1 < nr && L.off[lo + 1] <= e)) {
        int a = lo < 0 ? 0 : lo + 1, b = nr - 1;  // last staged row starting at or before e
        while (a < b) {
          const int mid = (a + b + 1) >> 1;
          if (L.off[mid] <= e) a = mid; else b = mid - 1;
        }
        lo = a;
        const uint32_t u = r0 + lo;
        fu = bit_set(nz_front, u) ? front[u] & lanes : 0;
      }
      if (fu) {
        w = cols[e0 + e];
        if (fu & ~__atomic_load_n(next + w, __ATOMIC_RELAXED))
          first = atomicOr((unsigned long long*)(next + w), (unsigned long long)fu) == 0;
      }
    }
    const uint64_t b = __ballot(first);
    if (b) {
      unsigned long long pos = 0;
      if (lane == 0) pos = atomicAdd(touch_count, (unsigned long long)__popcll(b));
      pos = __shfl(pos, 0, 64);
      if (first) touch[pos + __popcll(b & ((UINT64_C(1) << lane) - 1))] = w;
    }
  }
  __syncthreads();
}

// Chunks with a frontier row -> list (wave-aggregated appends to *count,
// which must be zero beforehand); order within the list is arbitrary.
__device__ __forceinline__ void ms_collect_chunks_body(const uint32_t* __restrict__ chunk_rows, uint64_t nchunks,
                                                       const uint64_t* __restrict__ nz_front,
                                                       uint32_t* __restrict__ list, unsigned* count, unsigned bid,
                                                       unsigned nblk) {
  const int lane = threadIdx.x & 63;
  const uint64_t stride = (uint64_t)nblk * kBlock;
  for (uint64_t base = (uint64_t)bid * kBlock + (threadIdx.x & ~63u); base < nchunks; base += stride) {
    const uint64_t c = base + lane;
    const bool act = c < nchunks && chunk_active(c, chunk_rows, nz_front);
    const uint64_t b = __ballot(act);
    if (!b) continue;
    unsigned pos = 0;
    if (lane == 0) pos = atomicAdd(count, (unsigned)__popcll(b));
    pos = __shfl(pos, 0, 64);
    if (act) list[pos + __popcll(b & ((UINT64_C(1) << lane) - 1))] = (uint32_t)c;
  }
}

// Push over the collected chunks: front[u] to every target w of a frontier
// row u, with a filtered 64-bit atomicOr (next zero beforehand).
__device__ __forceinline__ void ms_push_list_body(const uint32_t* __restrict__ list, const unsigned* count,
                                                  const uint32_t* __restrict__ chunk_rows, uint64_t nnz,
                                                  const uint64_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ cols,
                                                  const uint64_t* __restrict__ front,
                                                  const uint64_t* __restrict__ nz_front, uint64_t* next, SweepLds& L,
                                                  unsigned bid, unsigned nblk, uint64_t lanes = ~UINT64_C(0)) {
  const unsigned n = __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (unsigned i = bid; i < n; i += nblk)
    push_chunk(list[i], chunk_rows, nnz, offsets, cols, front, nz_front, next, L, lanes);
}

// One bit per vertex whose frontier word holds one of `lanes`.
__device__ __forceinline__ void ms_select_body(const uint64_t* __restrict__ front, int64_t rows, uint64_t lanes,
                                               uint64_t* __restrict__ nz_sel, unsigned bid, unsigned nblk) {
  const int lane = threadIdx.x & 63;
  const int64_t words = (rows + 63) / 64;
  const int64_t nwaves = (int64_t)nblk * (kBlock / 64);
  for (int64_t wd = (int64_t)bid * (kBlock / 64) + (threadIdx.x >> 6); wd < words; wd += nwaves) {
    const int64_t v = wd * 64 + lane;
    const uint64_t b = __ballot(v < rows && (front[v] & lanes) != 0);
    if (lane == 0) nz_sel[wd] = b;
  }
}

__device__ __forceinline__ void ms_zero_body(uint64_t* __restrict__ a, int64_t n, unsigned bid, unsigned nblk) {
  for (int64_t i = (int64_t)bid * kBlock + threadIdx.x; i < n; i += (int64_t)nblk * kBlock) a[i] = 0;
}


}  // namespace g5
