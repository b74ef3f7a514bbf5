// g5_frontier.hpp -- device helpers shared by the single-root and the batched
// BFS: frontier-degree scans, the chunked virtual edge list of a frontier and
// block reductions.
#pragma once

#include "g5_prims.hpp"

namespace g5 {

constexpr int kEpt = 4;              // frontier edges per lane and chunk
constexpr int kEpb = kEpt * kBlock;  // frontier edges per workgroup chunk

static inline unsigned grid_of(int64_t n, int64_t per, int64_t cap) {
  int64_t b = (n + per - 1) / per;
  if (b > cap) b = cap;
  return (unsigned)(b > 0 ? b : 1);
}


// prefix input: degree of frontier vertex i (0 past the end).
struct FrontierDegree {
  const uint32_t* queue;
  const uint64_t* offsets;
  int64_t nf;
  __device__ uint64_t operator()(size_t i) const {
    if ((int64_t)i >= nf) return 0;
    const uint32_t u = queue[i];
    return offsets[u + 1] - offsets[u];
  }
};

// The same with the frontier's length read on the device (a queue whose
// count a scan wrote there); the scan may run over an upper bound.
struct FrontierDegreeN {
  const uint32_t* queue;
  const uint64_t* offsets;
  const uint64_t* nf;
  __device__ uint64_t operator()(size_t i) const {
    if (i >= *nf) return 0;
    const uint32_t u = queue[i];
    return offsets[u + 1] - offsets[u];
  }
};

__device__ __forceinline__ bool bit_set(const uint64_t* __restrict__ bm, uint32_t v) {
  return (bm[v >> 6] >> (v & 63)) & 1;
}

// Block-wide sum added to *acc with one atomic per workgroup.
__device__ __forceinline__ void block_add(uint64_t v, unsigned long long* acc) {
  __shared__ uint64_t s_part[kBlock / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int i = 0; i < kBlock / 64; ++i) t += s_part[i];
    if (t) atomicAdd(acc, (unsigned long long)t);
  }
  __syncthreads();  // s_part may be reused by the next call
}


// For every kEpb-edge chunk of the virtual edge list, the first and last
// frontier index whose row overlaps it (binary searches over the prefix, one
// lane per chunk boundary, so their latency overlaps across the grid).
template <int kChunk>
__global__ void __launch_bounds__(kBlock) k_chunk_rows(const uint64_t* __restrict__ prefix, int64_t nf,
                                                       uint64_t nchunks, uint32_t* __restrict__ rows_out) {
  const uint64_t E = prefix[nf];
  // nchunks may be an upper bound (a launch made before E is known on the host)
  const uint64_t nc = (E + kChunk - 1) / kChunk < nchunks ? (E + kChunk - 1) / kChunk : nchunks;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < 2 * nc;
       t += (uint64_t)gridDim.x * kBlock) {
    const uint64_t c = t >> 1;
    uint64_t key = c * kChunk;
    if (t & 1) key = ((key + kChunk < E) ? key + kChunk : E) - 1;
    int64_t lo = 0, hi = nf - 1;  // last i with prefix[i] <= key
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= key) lo = mid; else hi = mid - 1;
    }
    rows_out[t] = (uint32_t)lo;
  }
}

// LDS view of the frontier rows overlapping one kEpb-edge chunk.
struct ChunkRows {
  uint64_t pre[kEpb + 1];   // first virtual edge index of each row
  uint64_t base[kEpb + 1];  // offsets[u] - pre: cols index = base + virtual index
  uint32_t src[kEpb + 1];   // the frontier vertex u
};

// Stages the rows of chunk c (every lane of the block must call it; it ends
// with a barrier) and returns their number.  Every frontier vertex has a
// non-empty row, so at most kEpb rows overlap a chunk; the clamp only keeps
// LDS indexing in bounds.
__device__ __forceinline__ int stage_chunk(uint64_t c, const uint32_t* __restrict__ chunk_rows,
                                           const uint32_t* __restrict__ queue,
                                           const uint64_t* __restrict__ prefix,
                                           const uint64_t* __restrict__ offsets, ChunkRows& cr) {
  const int64_t i0 = chunk_rows[2 * c];
  int cnt = (int)((int64_t)chunk_rows[2 * c + 1] - i0 + 1);
  cnt = cnt > kEpb + 1 ? kEpb + 1 : cnt;
  for (int j = threadIdx.x; j < cnt; j += kBlock) {
    const uint32_t u = queue[i0 + j];
    const uint64_t p = prefix[i0 + j];
    cr.pre[j] = p;
    cr.base[j] = offsets[u] - p;
    cr.src[j] = u;
  }
  __syncthreads();
  return cnt;
}

// Calls f(target, source) for this lane's kEpt consecutive edges of [e0, e1):
// one LDS binary search, then a forward walk across row boundaries.
template <class F>
__device__ __forceinline__ void walk_lane(uint64_t e0, uint64_t e1, int cnt, const ChunkRows& cr,
                                          const uint32_t* __restrict__ cols, F&& f) {
  const uint64_t my0 = e0 + (uint64_t)threadIdx.x * kEpt;
  if (my0 >= e1) return;
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cr.pre[mid] <= my0) lo = mid; else hi = mid - 1;
  }
  const uint64_t my1 = (my0 + kEpt < e1) ? my0 + kEpt : e1;
  for (uint64_t e = my0; e < my1; ++e) {
    while (lo + 1 < cnt && cr.pre[lo + 1] <= e) ++lo;
    f(cols[cr.base[lo] + e], cr.src[lo]);
  }
}

// Calls f(target, source) for the edges e0 + threadIdx.x + j * kBlock of
// [e0, e1): consecutive lanes take consecutive edges, so a wave reads 256
// contiguous bytes of cols per step and, within a sorted hub row, probes
// targets that share few cache lines.  A lane finds each edge's row by
// binary search past the row of its previous edge (none inside a long row).
template <class F>
__device__ __forceinline__ void walk_strided(uint64_t e0, uint64_t e1, int cnt, const ChunkRows& cr,
                                             const uint32_t* __restrict__ cols, F&& f) {
  int lo = 0;
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
    if (lo + 1 < cnt && cr.pre[lo + 1] <= e) {
      int a = lo + 1, b = cnt - 1;  // last row starting at or before e
      while (a < b) {
        const int mid = (a + b + 1) >> 1;
        if (cr.pre[mid] <= e) a = mid; else b = mid - 1;
      }
      lo = a;
    }
    f(cols[cr.base[lo] + e], cr.src[lo]);
  }
}

// The same edges as walk_strided, gathered first: a lane's (up to kEpt)
// targets and sources land in registers before any of them is used, so the
// column loads, and then the caller's probes, are in flight together.
__device__ __forceinline__ void gather_strided(uint64_t e0, uint64_t e1, int cnt, const ChunkRows& cr,
                                               const uint32_t* __restrict__ cols, uint32_t (&w)[kEpt],
                                               uint32_t (&u)[kEpt], bool (&ok)[kEpt]) {
  int lo = 0;
#pragma unroll
  for (int j = 0; j < kEpt; ++j) {
    const uint64_t e = e0 + threadIdx.x + (uint64_t)j * kBlock;
    ok[j] = e < e1;
    w[j] = 0;
    u[j] = 0;
    if (ok[j]) {
      if (lo + 1 < cnt && cr.pre[lo + 1] <= e) {
        int a = lo + 1, b = cnt - 1;
        while (a < b) {
          const int mid = (a + b + 1) >> 1;
          if (cr.pre[mid] <= e) a = mid; else b = mid - 1;
        }
        lo = a;
      }
      w[j] = cols[cr.base[lo] + e];
      u[j] = cr.src[lo];
    }
  }
}

// Tuple weight of a reached vertex (0xFF = unreached); the traversed-edge
// count of a run is the sum over all rows, halved.
struct ReachedWeight {
  const uint8_t* level;
  const uint32_t* weight;
  __device__ uint64_t operator()(size_t i) const { return level[i] != 0xFF ? weight[i] : 0; }
};

}  // namespace g5
