// g5_measure.hip -- calibration kernels for the profiler's traffic counters
// (include/g5_measure.h).
#include "g5_internal.hpp"
#include "g5_prims.hpp"

#include "../../include/g5_measure.h"

namespace g5 {

__global__ void __launch_bounds__(kBlock) k_measure_fill(uint64_t* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    a[i] = (uint64_t)i * UINT64_C(0x9E3779B97F4A7C15);
}

// splitmix64 finaliser: a bijective mix of the read index
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
  z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
  return z ^ (z >> 31);
}

// Each lane reads 8 words per step (8 independent loads in flight), at
// positions mix64(read index) mod n; the XOR of what a workgroup read goes
// to out[blockIdx.x].
__global__ void __launch_bounds__(kBlock) k_measure_gather(const uint64_t* __restrict__ a, int64_t n,
                                                            int64_t gathers, uint64_t* __restrict__ out) {
  constexpr int kU = 8;
  uint64_t acc = 0;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kU; base < gathers; base += stride * kU) {
    uint64_t x[kU];
#pragma unroll
    for (int j = 0; j < kU; ++j) {
      const int64_t i = base + j;
      x[j] = i < gathers ? a[mix64((uint64_t)i) % (uint64_t)n] : 0;
    }
#pragma unroll
    for (int j = 0; j < kU; ++j) acc ^= x[j];
  }
  __shared__ uint64_t s[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc ^= __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t ^= s[w];
    out[blockIdx.x] = t;
  }
}

}  // namespace g5

using namespace g5;

extern "C" int g5_measure_gather(int device, int64_t array_bytes, int64_t gathers, double* ms, uint64_t* checksum) {
  if (array_bytes < 8 || array_bytes % 8 || gathers < 1 || !ms || !checksum) return G5_ERR_ARG;
  try {
    G5_HIP(hipSetDevice(device));
    const int64_t n = array_bytes / 8;
    const unsigned grid = 8192;
    DevBuf<uint64_t> a, out;
    a.alloc(n);
    out.alloc(grid);
    hipStream_t s = nullptr;
    G5_HIP(hipStreamCreate(&s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    G5_HIP(hipEventCreate(&e0));
    G5_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_measure_fill, dim3(grid), dim3(kBlock), 0, s, a.p, n);
    G5_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_measure_gather, dim3(grid), dim3(kBlock), 0, s, a.p, n, gathers, out.p);
    G5_HIP(hipGetLastError());
    G5_HIP(hipEventRecord(e1, s));
    std::vector<uint64_t> h(grid);
    G5_HIP(hipMemcpyAsync(h.data(), out.p, grid * 8, hipMemcpyDeviceToHost, s));
    G5_HIP(hipStreamSynchronize(s));
    float t = 0;
    G5_HIP(hipEventElapsedTime(&t, e0, e1));
    uint64_t x = 0;
    for (uint64_t v : h) x ^= v;
    *ms = t;
    *checksum = x;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    return G5_OK;
  } catch (const Error& e) {
    return e.code;
  }
}
