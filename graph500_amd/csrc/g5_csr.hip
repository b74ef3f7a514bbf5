// g5_csr.hip -- CSR construction in HBM.
//
// Restates the semantics of GraphConstructor2DCSR::construct
// (mpi/graph_constructor.hpp:554-574) for one GPU:
//   searchMaxVertex      estimated SCALE = msb(OR of all ids) + 1      (:735-763)
//   scatterAndScanEdges  self loops dropped; degree counts duplicates  (:765-843)
//   calc_degree VR=2     vertices ordered by degree, descending        (:296-352)
//                        (ties by original id: stable sort, the
//                        reference's std::sort leaves them unspecified)
//   sortEdges            rows sorted by reordered target, duplicates removed (:1376-1442)
//   isolateFirstEdge     first (smallest new id) neighbour kept apart  (:845-879)
//   computeNumVertices   n_active = vertices with an edge              (:1444-1464)
//
// Device pipeline: OR-reduce -> atomic degree count -> stable radix sort of
// (degree key, id) -> one uint64 key (new_src << S | new_tgt) per directed
// entry -> LSD radix sort over 2S bits -> one scan that drops duplicates,
// writes columns and row offsets.  The tuples themselves are never re-sorted.
#include "g5_ctx.hpp"
#include "g5_prims.hpp"

#include <chrono>

namespace g5 {

__device__ __forceinline__ uint64_t tuple_v0(const g5_packed_edge& e) {
  return (uint64_t)e.v0_low | ((uint64_t)(e.high & 0xFFFF) << 32);
}
__device__ __forceinline__ uint64_t tuple_v1(const g5_packed_edge& e) {
  return (uint64_t)e.v1_low | ((uint64_t)(e.high >> 16) << 32);
}

__global__ void __launch_bounds__(kBlock) k_or_ids(const g5_packed_edge* __restrict__ e, int64_t m,
                                                   unsigned long long* out) {
  uint64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    const g5_packed_edge t = e[i];
    acc |= tuple_v0(t) | tuple_v1(t);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc |= __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0 && acc) atomicOr(out, (unsigned long long)acc);
}

__global__ void __launch_bounds__(kBlock) k_degree(const g5_packed_edge* __restrict__ e, int64_t m,
                                                   uint32_t* __restrict__ deg,
                                                   uint32_t* __restrict__ loops) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    const g5_packed_edge t = e[i];
    const uint64_t a = tuple_v0(t), b = tuple_v1(t);
    if (a == b) {
      atomicAdd(&loops[a], 1u);
    } else {
      atomicAdd(&deg[a], 1u);
      atomicAdd(&deg[b], 1u);
    }
  }
}

// Sort keys for the vertex order: smaller key first, stable by id.
__global__ void k_order_keys(const uint32_t* __restrict__ deg, int64_t n, uint32_t maxdeg, int mode,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t d = deg[v];
    keys[v] = (mode == 2) ? (maxdeg - d) : (d == 0 ? 1u : 0u);
    ids[v] = (uint32_t)v;
  }
}

struct NonZeroU32 {
  const uint32_t* p;
  __device__ uint64_t operator()(size_t i) const { return p[i] ? 1 : 0; }
};

__global__ void k_iota(int64_t n, uint32_t* __restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    out[v] = (uint32_t)v;
}

__global__ void k_invert(const uint32_t* __restrict__ new2old, int64_t n, uint32_t* __restrict__ old2new) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    old2new[new2old[v]] = (uint32_t)v;
}

__global__ void k_weights(const uint32_t* __restrict__ new2old, const uint32_t* __restrict__ deg,
                          const uint32_t* __restrict__ loops, int64_t rows, uint32_t* __restrict__ w) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < rows;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t o = new2old[v];
    w[v] = deg[o] + 2u * loops[o];
  }
}

// One key per directed entry: (new src << S) | new tgt; loops get an all-ones
// key that sorts after every real key and is cut off by the count.
__global__ void __launch_bounds__(kBlock) k_pair_keys(const g5_packed_edge* __restrict__ e, int64_t m,
                                                      const uint32_t* __restrict__ old2new, int scale,
                                                      uint64_t* __restrict__ keys) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    const g5_packed_edge t = e[i];
    const uint64_t a = tuple_v0(t), b = tuple_v1(t);
    uint64_t k0 = ~UINT64_C(0), k1 = ~UINT64_C(0);
    if (a != b) {
      const uint64_t na = old2new[a], nb = old2new[b];
      k0 = (na << scale) | nb;
      k1 = (nb << scale) | na;
    }
    keys[2 * i] = k0;
    keys[2 * i + 1] = k1;
  }
}

struct UniqueFlag {
  const uint64_t* keys;
  size_t nvalid;
  __device__ uint64_t operator()(size_t i) const {
    if (i >= nvalid) return 0;
    return (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
  }
};

// Writes the column of every first occurrence and the row offsets of every
// row that starts at it (rows with no entries in between get the same start).
struct UniqueStore {
  const uint64_t* keys;
  size_t nvalid;
  int scale;
  uint32_t* cols;
  uint64_t* offsets;
  __device__ void operator()(size_t i, uint64_t pre) const {
    if (i >= nvalid) return;
    const uint64_t k = keys[i];
    if (i != 0 && k == keys[i - 1]) return;
    const uint64_t mask = (UINT64_C(1) << scale) - 1;
    cols[pre] = (uint32_t)(k & mask);
    const int64_t src = (int64_t)(k >> scale);
    const int64_t prev = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> scale);
    for (int64_t r = prev + 1; r <= src; ++r) offsets[r] = pre;
  }
};

__global__ void k_offsets_tail(const uint64_t* __restrict__ keys, size_t nvalid, int scale,
                               int64_t rows, uint64_t nnz, uint64_t* __restrict__ offsets) {
  const int64_t last = nvalid ? (int64_t)(keys[nvalid - 1] >> scale) : -1;
  for (int64_t r = last + 1 + (int64_t)(blockIdx.x * blockDim.x + threadIdx.x); r <= rows;
       r += (int64_t)gridDim.x * blockDim.x)
    offsets[r] = nnz;
}

__global__ void k_first_nbr(const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ cols,
                            int64_t rows, uint32_t* __restrict__ first) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < rows;
       v += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t b = offsets[v], e = offsets[v + 1];
    first[v] = (e > b) ? cols[b] : 0xFFFFFFFFu;
  }
}

// has_edge bitmap by original id (one wave writes one 64-bit word).
__global__ void __launch_bounds__(kBlock) k_has_edge_bits(const uint32_t* __restrict__ deg, int64_t n,
                                                          uint64_t* __restrict__ words) {
  for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v - (threadIdx.x & 63) < n;
       v += (int64_t)gridDim.x * kBlock) {
    const bool has = v < n && deg[v] > 0;
    const uint64_t w = __ballot(has);
    if ((threadIdx.x & 63) == 0) words[v >> 6] = w;
  }
}

static unsigned grid_for(int64_t n, int64_t per_block = kBlock) {
  int64_t b = (n + per_block - 1) / per_block;
  if (b > 65536) b = 65536;
  return (unsigned)(b > 0 ? b : 1);
}

void build_graph(g5_ctx* ctx, int mode, bool release_edges) {
  G5_CHECK(mode >= 0 && mode <= 2, "g5_construct: reorder_mode must be 0, 1 or 2");
  G5_CHECK(ctx->num_edges > 0, "g5_construct: the edge list is empty");
  hipStream_t s = ctx->stream;
  Graph& g = ctx->graph;
  g.clear();
  ctx->acc.alloc(4);
  const int64_t m = ctx->num_edges;
  const g5_packed_edge* e = ctx->edges.p;

  hipEvent_t t0, t1;
  G5_HIP(hipEventCreate(&t0));
  G5_HIP(hipEventCreate(&t1));
  G5_HIP(hipEventRecord(t0, s));

  // 1. estimated SCALE
  G5_HIP(hipMemsetAsync(ctx->acc.p, 0, 8, s));
  hipLaunchKernelGGL(k_or_ids, dim3(grid_for(m, kBlock * 8)), dim3(kBlock), 0, s, e, m,
                     (unsigned long long*)ctx->acc.p);
  G5_HIP(hipGetLastError());
  uint64_t orv = 0;
  G5_HIP(hipMemcpyAsync(&orv, ctx->acc.p, 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  int scale = 0;
  while (scale < 64 && (orv >> scale)) ++scale;
  G5_CHECK(scale >= 1, "g5_construct: all vertex ids are zero");
  if (scale > 32) throw Error{"g5_construct: vertex ids need " + std::to_string(scale) +
                              " bits; this build supports up to 2^32 vertices"};
  const int64_t n = (int64_t)1 << scale;
  g.scale = scale;
  g.n = n;
  g.num_tuples = m;
  g.reorder_mode = mode;

  // 2. degrees (duplicates counted, loops apart)
  DevBuf<uint32_t> deg, loops;
  deg.alloc(n);
  loops.alloc(n);
  G5_HIP(hipMemsetAsync(deg.p, 0, n * 4, s));
  G5_HIP(hipMemsetAsync(loops.p, 0, n * 4, s));
  hipLaunchKernelGGL(k_degree, dim3(grid_for(m, kBlock * 4)), dim3(kBlock), 0, s, e, m, deg.p, loops.p);
  G5_HIP(hipGetLastError());
  g.max_degree = (int64_t)device_reduce_max(LoadU32{deg.p}, (size_t)n, ctx->acc.p, s);
  g.self_loops = (int64_t)device_reduce_sum(LoadU32{loops.p}, (size_t)n, ctx->acc.p, s);
  g.n_active = (int64_t)device_reduce_sum(NonZeroU32{deg.p}, (size_t)n, ctx->acc.p, s);
  g.directed_before = 2 * (m - g.self_loops);
  g.n_rows = (mode == 0) ? n : g.n_active;

  // has_edge bitmap by original id, kept on the host for root selection
  {
    DevBuf<uint64_t> words;
    const int64_t nw = (n + 63) / 64;
    words.alloc(nw);
    hipLaunchKernelGGL(k_has_edge_bits, dim3(grid_for(n)), dim3(kBlock), 0, s, deg.p, n, words.p);
    G5_HIP(hipGetLastError());
    g.has_edge_words.resize(nw);
    G5_HIP(hipMemcpyAsync(g.has_edge_words.data(), words.p, nw * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    g.max_used_vertex = 0;
    for (int64_t w = nw - 1; w >= 0; --w)
      if (g.has_edge_words[w]) {
        g.max_used_vertex = w * 64 + 63 - __builtin_clzll(g.has_edge_words[w]);
        break;
      }
  }

  // 3. vertex order
  g.new2old.alloc(n);
  g.old2new.alloc(n);
  RadixScratch rs;
  if (mode == 0) {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, n, g.new2old.p);
    G5_HIP(hipGetLastError());
    G5_HIP(hipMemcpyAsync(g.old2new.p, g.new2old.p, n * 4, hipMemcpyDeviceToDevice, s));
  } else {
    DevBuf<uint32_t> k0, k1, v1;
    k0.alloc(n);
    k1.alloc(n);
    v1.alloc(n);
    hipLaunchKernelGGL(k_order_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, deg.p, n,
                       (uint32_t)g.max_degree, mode, k0.p, g.new2old.p);
    G5_HIP(hipGetLastError());
    int bits = 0;
    const uint64_t kmax = (mode == 2) ? (uint64_t)g.max_degree : 1;
    while (bits < 32 && (kmax >> bits)) ++bits;
    uint32_t *kp = k0.p, *ka = k1.p, *vp = g.new2old.p, *va = v1.p;
    radix_sort_pairs_u32(&kp, &ka, &vp, &va, (size_t)n, 0, bits, rs, s);
    if (vp != g.new2old.p)
      G5_HIP(hipMemcpyAsync(g.new2old.p, vp, n * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_invert, dim3(grid_for(n)), dim3(kBlock), 0, s, g.new2old.p, n, g.old2new.p);
    G5_HIP(hipGetLastError());
    G5_HIP(hipStreamSynchronize(s));
  }
  const int64_t rows = g.n_rows;
  g.weight.alloc(rows);
  hipLaunchKernelGGL(k_weights, dim3(grid_for(rows)), dim3(kBlock), 0, s, g.new2old.p, deg.p, loops.p,
                     rows, g.weight.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipStreamSynchronize(s));
  deg.release();
  loops.release();

  // 4. directed entry keys and their sort
  const size_t nkeys = 2 * (size_t)m;
  const size_t nvalid = (size_t)g.directed_before;
  DevBuf<uint64_t> keys, alt;
  keys.alloc(nkeys);
  hipLaunchKernelGGL(k_pair_keys, dim3(grid_for(m, kBlock * 4)), dim3(kBlock), 0, s, e, m,
                     g.old2new.p, scale, keys.p);
  G5_HIP(hipGetLastError());
  if (release_edges) {
    G5_HIP(hipStreamSynchronize(s));
    ctx->edges.release();
    ctx->num_edges = 0;
  }
  alt.alloc(nkeys);
  uint64_t *kp = keys.p, *ka = alt.p;
  radix_sort_u64(&kp, &ka, nkeys, 0, 2 * scale, rs, s);
  G5_HIP(hipStreamSynchronize(s));
  rs.counts.release();
  rs.offsets.release();
  // keep the sorted buffer, free the other
  if (kp == keys.p) alt.release(); else keys.release();
  const uint64_t* sorted = kp;

  // 5. duplicate removal, columns, offsets
  const uint64_t nnz = device_reduce_sum(UniqueFlag{sorted, nvalid}, nkeys, ctx->acc.p, s);
  g.nnz = (int64_t)nnz;
  g.cols.alloc(nnz);
  g.offsets.alloc(rows + 1);
  DevBuf<uint64_t> tmp;
  tmp.alloc(scan_tmp_elems(nvalid));
  device_scan(UniqueFlag{sorted, nvalid}, UniqueStore{sorted, nvalid, scale, g.cols.p, g.offsets.p},
              nvalid, tmp.p, s);
  hipLaunchKernelGGL(k_offsets_tail, dim3(grid_for(rows)), dim3(kBlock), 0, s, sorted, nvalid, scale,
                     rows, nnz, g.offsets.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipStreamSynchronize(s));
  keys.release();
  alt.release();
  tmp.release();

  g.first_nbr.alloc(rows);
  hipLaunchKernelGGL(k_first_nbr, dim3(grid_for(rows)), dim3(kBlock), 0, s, g.offsets.p, g.cols.p,
                     rows, g.first_nbr.p);
  G5_HIP(hipGetLastError());
  G5_HIP(hipEventRecord(t1, s));
  G5_HIP(hipEventSynchronize(t1));
  float ms = 0;
  G5_HIP(hipEventElapsedTime(&ms, t0, t1));
  g.construct_ms = ms;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  g.built = true;
}

}  // namespace g5
