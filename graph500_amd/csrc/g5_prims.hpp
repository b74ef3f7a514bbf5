// g5_prims.hpp -- device-wide building blocks (scan, reduce, stable LSD radix
// sort) written for wave64 / 256-thread workgroups on gfx950.
#pragma once

#include "g5_internal.hpp"

namespace g5 {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;  // 4096 elements per workgroup tile

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane ? (~UINT64_C(0) >> (64 - lane)) : 0;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
// 'tot' receives the block total.  Needs 4+1 words of LDS.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* s_wave, uint64_t& tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = wave_incl_scan(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint64_t off = 0;
  uint64_t all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const uint64_t x = s_wave[w];
    off += (w < wave) ? x : 0;
    all += x;
  }
  __syncthreads();
  tot = all;
  return off + inc - v;
}

// ---------------------------------------------------------------------------
// Scan: in(i) -> uint64 values; out(i, exclusive_prefix, value).
// ---------------------------------------------------------------------------
template <class In>
__global__ void __launch_bounds__(kBlock) k_scan_partials(In in, size_t n, uint64_t* partial) {
  __shared__ uint64_t s_wave[kBlock / 64];
  const size_t base = (size_t)blockIdx.x * kTile;
  uint64_t sum = 0;
#pragma unroll 4
  for (int j = 0; j < kItems; ++j) {
    const size_t i = base + (size_t)j * kBlock + threadIdx.x;
    if (i < n) sum += in(i);
  }
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_wave[w];
    partial[blockIdx.x] = t;
  }
}

template <class In, class Out>
__global__ void __launch_bounds__(kBlock) k_scan_final(In in, Out out, size_t n,
                                                       const uint64_t* block_off) {
  __shared__ uint64_t s_tile[kTile];
  __shared__ uint64_t s_wave[kBlock / 64];
  const size_t base = (size_t)blockIdx.x * kTile;
  for (int j = 0; j < kItems; ++j) {
    const size_t i = base + (size_t)j * kBlock + threadIdx.x;
    s_tile[j * kBlock + threadIdx.x] = (i < n) ? in(i) : 0;
  }
  __syncthreads();
  uint64_t local[kItems];
  uint64_t sum = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    local[j] = s_tile[threadIdx.x * kItems + j];
    sum += local[j];
  }
  uint64_t tot;
  uint64_t run = block_excl_scan(sum, s_wave, tot) + (block_off ? block_off[blockIdx.x] : 0);
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    s_tile[threadIdx.x * kItems + j] = run;
    run += local[j];
  }
  __syncthreads();
  for (int j = 0; j < kItems; ++j) {
    const int p = j * kBlock + threadIdx.x;
    const size_t i = base + p;
    if (i < n) {
      // the value is recovered from the neighbouring prefix
      const uint64_t pre = s_tile[p];
      out(i, pre);
    }
  }
}

struct LoadU64 {
  const uint64_t* p;
  __device__ uint64_t operator()(size_t i) const { return p[i]; }
};
struct LoadU32 {
  const uint32_t* p;
  __device__ uint64_t operator()(size_t i) const { return p[i]; }
};
struct StoreU64 {
  uint64_t* p;
  __device__ void operator()(size_t i, uint64_t pre) const { p[i] = pre; }
};

// Exclusive scan of 'partial' (nb values) in place, recursively.
void scan_partials_inplace(uint64_t* partial, size_t nb, uint64_t* tmp, hipStream_t s);

// Generic exclusive scan driver.  tmp must hold scan_tmp_elems(n) uint64.
template <class In, class Out>
void device_scan(In in, Out out, size_t n, uint64_t* tmp, hipStream_t s) {
  if (n == 0) return;
  const size_t nb = (n + kTile - 1) / kTile;
  uint64_t* partial = tmp;
  if (nb > 1) {
    hipLaunchKernelGGL(k_scan_partials<In>, dim3((unsigned)nb), dim3(kBlock), 0, s, in, n, partial);
    G5_HIP(hipGetLastError());
    scan_partials_inplace(partial, nb, tmp + nb, s);
  }
  hipLaunchKernelGGL((k_scan_final<In, Out>), dim3((unsigned)nb), dim3(kBlock), 0, s, in, out, n,
                     nb > 1 ? partial : (const uint64_t*)nullptr);
  G5_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Reduction: sum of in(i) over [0, n) into *d_out (atomic, d_out pre-zeroed).
// ---------------------------------------------------------------------------
template <class In>
__global__ void __launch_bounds__(kBlock) k_reduce_sum(In in, size_t n, unsigned long long* d_out) {
  __shared__ uint64_t s_wave[kBlock / 64];
  uint64_t sum = 0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    sum += in(i);
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += s_wave[w];
    if (t) atomicAdd(d_out, (unsigned long long)t);
  }
}

template <class In>
__global__ void __launch_bounds__(kBlock) k_reduce_max(In in, size_t n, unsigned long long* d_out) {
  __shared__ uint64_t s_wave[kBlock / 64];
  uint64_t m = 0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    const uint64_t v = in(i);
    m = v > m ? v : m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t t = __shfl_xor(m, o, 64);
    m = t > m ? t : m;
  }
  if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t = s_wave[w] > t ? s_wave[w] : t;
    atomicMax(d_out, (unsigned long long)t);
  }
}

inline unsigned reduce_grid(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 2048) b = 2048;
  return (unsigned)(b ? b : 1);
}

// Sum / max of in(i) returned to the host (synchronizes the stream).
template <class In>
uint64_t device_reduce_sum(In in, size_t n, uint64_t* d_acc, hipStream_t s) {
  G5_HIP(hipMemsetAsync(d_acc, 0, sizeof(uint64_t), s));
  if (n) {
    hipLaunchKernelGGL(k_reduce_sum<In>, dim3(reduce_grid(n)), dim3(kBlock), 0, s, in, n,
                       (unsigned long long*)d_acc);
    G5_HIP(hipGetLastError());
  }
  uint64_t h = 0;
  G5_HIP(hipMemcpyAsync(&h, d_acc, sizeof(h), hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  return h;
}

template <class In>
uint64_t device_reduce_max(In in, size_t n, uint64_t* d_acc, hipStream_t s) {
  G5_HIP(hipMemsetAsync(d_acc, 0, sizeof(uint64_t), s));
  if (n) {
    hipLaunchKernelGGL(k_reduce_max<In>, dim3(reduce_grid(n)), dim3(kBlock), 0, s, in, n,
                       (unsigned long long*)d_acc);
    G5_HIP(hipGetLastError());
  }
  uint64_t h = 0;
  G5_HIP(hipMemcpyAsync(&h, d_acc, sizeof(h), hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace g5
