// g5_prims.hip -- scan and stable LSD radix sort for gfx950.
//
// Radix sort, per 8-bit digit pass (3 launches):
//   1. k_radix_hist     per-tile digit histogram -> counts[digit][tile] (uint32)
//   2. device scan      counts -> global output start of (digit, tile) (uint64)
//   3. k_radix_scatter  stable in-tile ranking with wave64 ballots: digits are
//                       matched across the wave with 8 ballots, waves are
//                       ordered through an LDS table, tiles through the scan;
//                       keys are first placed digit-sorted in LDS, then
//                       written out so each digit run is a contiguous store.
// Tiles are 4096 elements (256 lanes x 16), read in striped order so every
// global load is coalesced.
#include "g5_prims.hpp"

namespace g5 {

size_t scan_tmp_elems(size_t n) {
  size_t total = 0;
  size_t nb = (n + kTile - 1) / kTile;
  while (nb > 1) {
    total += nb;
    nb = (nb + kTile - 1) / kTile;
  }
  return total + 2;
}

void scan_partials_inplace(uint64_t* partial, size_t nb, uint64_t* tmp, hipStream_t s) {
  device_scan(LoadU64{partial}, StoreU64{partial}, nb, tmp, s);
}

uint64_t scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, uint64_t* tmp,
                         hipStream_t s, bool want_total) {
  device_scan(LoadU32{in}, StoreU64{out}, n, tmp, s);
  if (!want_total || n == 0) return 0;
  uint64_t last_pre = 0;
  uint32_t last = 0;
  G5_HIP(hipMemcpyAsync(&last_pre, out + n - 1, 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipMemcpyAsync(&last, in + n - 1, 4, hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  return last_pre + last;
}

uint64_t scan_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t* tmp, hipStream_t s,
                  bool want_total) {
  uint64_t last = 0;
  if (want_total && n) {
    G5_HIP(hipMemcpyAsync(&last, in + n - 1, 8, hipMemcpyDeviceToHost, s));
  }
  device_scan(LoadU64{in}, StoreU64{out}, n, tmp, s);
  if (!want_total || n == 0) return 0;
  uint64_t last_pre = 0;
  G5_HIP(hipMemcpyAsync(&last_pre, out + n - 1, 8, hipMemcpyDeviceToHost, s));
  G5_HIP(hipStreamSynchronize(s));
  return last_pre + last;
}

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
template <class K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
  return (uint32_t)((k >> shift) & 0xFF);
}

template <class K>
__global__ void __launch_bounds__(kBlock) k_radix_hist(const K* __restrict__ keys, size_t n,
                                                       int shift, uint32_t* __restrict__ counts,
                                                       size_t num_tiles) {
  __shared__ uint32_t s_hist[256];
  s_hist[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int j = 0; j < kItems; ++j) {
    const size_t i = base + (size_t)j * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&s_hist[digit_of(keys[i], shift)], 1u);
  }
  __syncthreads();
  counts[(size_t)threadIdx.x * num_tiles + blockIdx.x] = s_hist[threadIdx.x];
}

template <class K, bool kVals>
__global__ void __launch_bounds__(kBlock) k_radix_scatter(
    const K* __restrict__ keys_in, K* __restrict__ keys_out, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ vals_out, size_t n, int shift, const uint32_t* __restrict__ counts,
    const uint64_t* __restrict__ offsets, size_t num_tiles) {
  __shared__ K s_keys[kTile];
  __shared__ uint32_t s_vals[kVals ? kTile : 1];
  __shared__ uint32_t s_hist[kBlock / 64][256];
  __shared__ uint32_t s_run[256];
  __shared__ uint32_t s_start[256];
  __shared__ uint64_t s_global[256];
  __shared__ uint64_t s_wave[kBlock / 64];

  const int t = threadIdx.x, wave = t >> 6;
  const size_t tile = blockIdx.x;
  const size_t base = tile * kTile;

  // Digit t: where its keys start inside this tile and in the output.
  const uint32_t my_count = counts[(size_t)t * num_tiles + tile];
  s_global[t] = offsets[(size_t)t * num_tiles + tile];
  uint64_t tot;
  const uint64_t start = block_excl_scan(my_count, s_wave, tot);
  s_start[t] = (uint32_t)start;
  s_run[t] = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) s_hist[w][t] = 0;
  __syncthreads();

  const uint64_t lt = lanemask_lt();
  for (int r = 0; r < kItems; ++r) {
    const size_t i = base + (size_t)r * kBlock + t;
    const bool valid = i < n;
    K k = valid ? keys_in[i] : K(0);
    uint32_t v = 0;
    if (kVals && valid) v = vals_in[i];
    const uint32_t d = digit_of(k, shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const bool leader = valid && ((peers & lt) == 0);
    if (leader) s_hist[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_start[d] + s_run[d] + rank;
#pragma unroll
      for (int w = 0; w < kBlock / 64; ++w) pos += (w < wave) ? s_hist[w][d] : 0;
      s_keys[pos] = k;
      if (kVals) s_vals[pos] = v;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
      add += s_hist[w][t];
      s_hist[w][t] = 0;
    }
    s_run[t] += add;
    __syncthreads();
  }

  const size_t cnt = (n - base) < (size_t)kTile ? (n - base) : (size_t)kTile;
  for (int r = 0; r < kItems; ++r) {
    const int p = r * kBlock + t;
    if ((size_t)p < cnt) {
      const K k = s_keys[p];
      const uint32_t d = digit_of(k, shift);
      const uint64_t o = s_global[d] + (uint64_t)(p - (int)s_start[d]);
      keys_out[o] = k;
      if (kVals) vals_out[o] = s_vals[p];
    }
  }
}

template <class K, bool kVals>
static void radix_sort_impl(K** keys, K** alt, uint32_t** vals, uint32_t** alt_vals, size_t n,
                            int begin_bit, int end_bit, RadixScratch& sc, hipStream_t s) {
  if (n <= 1 || end_bit <= begin_bit) return;
  const size_t tiles = (n + kTile - 1) / kTile;
  const size_t ncounts = tiles * 256;
  sc.counts.alloc(ncounts);
  sc.offsets.alloc(ncounts);
  sc.scan_tmp.alloc(scan_tmp_elems(ncounts));
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist<K>, dim3((unsigned)tiles), dim3(kBlock), 0, s, *keys, n, shift,
                       sc.counts.p, tiles);
    G5_HIP(hipGetLastError());
    scan_u32_to_u64(sc.counts.p, sc.offsets.p, ncounts, sc.scan_tmp.p, s, false);
    hipLaunchKernelGGL((k_radix_scatter<K, kVals>), dim3((unsigned)tiles), dim3(kBlock), 0, s, *keys,
                       *alt, kVals ? *vals : nullptr, kVals ? *alt_vals : nullptr, n, shift,
                       sc.counts.p, sc.offsets.p, tiles);
    G5_HIP(hipGetLastError());
    std::swap(*keys, *alt);
    if (kVals) std::swap(*vals, *alt_vals);
  }
}

void radix_sort_u64(uint64_t** keys, uint64_t** alt, size_t n, int begin_bit, int end_bit,
                    RadixScratch& sc, hipStream_t s) {
  radix_sort_impl<uint64_t, false>(keys, alt, nullptr, nullptr, n, begin_bit, end_bit, sc, s);
}

void radix_sort_pairs_u32(uint32_t** keys, uint32_t** alt_keys, uint32_t** vals,
                          uint32_t** alt_vals, size_t n, int begin_bit, int end_bit,
                          RadixScratch& sc, hipStream_t s) {
  radix_sort_impl<uint32_t, true>(keys, alt_keys, vals, alt_vals, n, begin_bit, end_bit, sc, s);
}

}  // namespace g5
