// g5_internal.hpp -- shared definitions of the MI355X Graph500 hot path.
//
// Device memory layout (all arrays in HBM, sized for one graph):
//   edge list   : g5_packed_edge[m]            12-byte tuples (reference layout)
//   CSR         : uint64 offsets[n_active+1], uint32 cols[nnz], rows and columns
//                 in degree order (new ids), each row ascending, deduplicated
//   first_nbr   : uint32[n_active]             cols[offsets[v]] (highest-degree
//                 neighbour), read coalesced by the bottom-up sweep
//   weight      : uint32[n_active]             tuple endpoints per vertex
//                 (non-loop degree + 2 * self loops) for the TEPS count
//   old2new, new2old : uint32[n]
//   BFS state   : visited bitmaps uint64[2][W], level uint8[n_active],
//                 parent uint32[n_active], frontier queues uint32[2][n_active]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "../../include/g5.h"

namespace g5 {

struct Error {
  std::string msg;
  int code = G5_ERR_ARG;
};

#define G5_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::g5::Error{std::string(#call) + ": " + hipGetErrorString(e_) + " (" +    \
                        __FILE__ + ":" + std::to_string(__LINE__) + ")", G5_ERR_HIP}; \
  } while (0)

#define G5_CHECK(cond, text)                        \
  do {                                              \
    if (!(cond)) throw ::g5::Error{std::string(text)}; \
  } while (0)

#define G5_CHECK_CODE(cond, code, text)                       \
  do {                                                        \
    if (!(cond)) throw ::g5::Error{std::string(text), (code)}; \
  } while (0)

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Device buffer with explicit ownership (hipMalloc / hipFree).
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error{"hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                      hipGetErrorString(e),
                  G5_ERR_MEM};
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// Generic device primitives (g5_prims.hip)
// ---------------------------------------------------------------------------
// Exclusive scan of 'n' uint32 counts into uint64 offsets; returns the total.
// 'out' may not alias 'in'. Uses 'tmp' scratch sized by scan_tmp_elems(n).
size_t scan_tmp_elems(size_t n);
uint64_t scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, uint64_t* tmp,
                         hipStream_t s, bool want_total);
// Exclusive scan of uint64 in place-capable (in may equal out).
uint64_t scan_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t* tmp, hipStream_t s,
                  bool want_total);

// Stable LSD radix sort.  Keys-only uint64 over bits [begin_bit, end_bit).
// Result ends in *keys (buffers are swapped as needed).
struct RadixScratch {
  DevBuf<uint32_t> counts;
  DevBuf<uint64_t> offsets;
  DevBuf<uint64_t> scan_tmp;
};
void radix_sort_u64(uint64_t** keys, uint64_t** alt, size_t n, int begin_bit, int end_bit,
                    RadixScratch& scratch, hipStream_t s);
// Stable LSD radix sort of (uint32 key, uint32 value) pairs over [begin_bit, end_bit).
void radix_sort_pairs_u32(uint32_t** keys, uint32_t** alt_keys, uint32_t** vals,
                          uint32_t** alt_vals, size_t n, int begin_bit, int end_bit,
                          RadixScratch& scratch, hipStream_t s);

// ---------------------------------------------------------------------------
// MRG tables (host side, g5_gen.hip)
// ---------------------------------------------------------------------------
struct MrgEntry {
  uint32_t s, t, u, v, w, a, b, c, d;
};
const MrgEntry* host_mrg_table();  // [24][256]

struct MrgState {
  uint32_t z1, z2, z3, z4, z5;
};
void host_make_mrg_seed(uint64_t s1, uint64_t s2, MrgState* st);
void host_mrg_skip(MrgState* st, uint64_t hi, uint64_t mid, uint64_t lo);
uint32_t host_mrg_get_uint(MrgState* st);
double host_mrg_get_double(MrgState* st);

}  // namespace g5
