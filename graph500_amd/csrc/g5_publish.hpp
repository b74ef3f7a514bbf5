// g5_publish.hpp -- values a kernel hands to the host through coherent
// (pinned, fine-grained) host memory, with no copy command on the stream.
//
// A kernel's stores to pinned host memory need not become visible to the
// host in the order the kernel issued them: a value stored before a
// system-scope fence has been read stale on the host while values stored
// after the fence, in another 64-byte line, were already fresh.  So no value
// here is guarded by a separate flag: every value travels with the sequence
// number of the launch that publishes it, in one naturally aligned 16-byte
// store, and the host takes a value only once that value's own tag carries
// the sequence number it waits for.
#pragma once

#include "g5_internal.hpp"

namespace g5 {

struct alignas(16) Published {
  uint64_t value;
  uint64_t seq;
};

// One 16-byte store: value and tag become visible together.
__device__ __forceinline__ void publish(Published* p, uint64_t value, uint64_t seq) {
  typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
  *reinterpret_cast<u64x2*>(p) = u64x2{(unsigned long long)value, (unsigned long long)seq};
}

// A block of n published values in coherent host memory.  Sequence numbers
// start at 1; the block starts with every tag 0.
struct PublishBlock {
  Published* h = nullptr;  // host address
  Published* d = nullptr;  // the same block as kernels address it
  int n = 0;

  PublishBlock() = default;
  PublishBlock(const PublishBlock&) = delete;
  PublishBlock& operator=(const PublishBlock&) = delete;
  ~PublishBlock() { release(); }

  void alloc(int count) {
    release();
    G5_HIP(hipHostMalloc((void**)&h, (size_t)count * sizeof(Published),
                         hipHostMallocCoherent | hipHostMallocMapped));
    for (int i = 0; i < count; ++i) h[i] = Published{0, 0};
    G5_HIP(hipHostGetDevicePointer((void**)&d, h, 0));
    n = count;
  }
  void release() {
    if (h) (void)hipHostFree(h);
    h = d = nullptr;
    n = 0;
  }
  // true once every value of [i0, i1) carries seq
  bool ready(int i0, int i1, uint64_t seq) const {
    const volatile Published* v = h;
    for (int i = i0; i < i1; ++i)
      if (v[i].seq != seq) return false;
    return true;
  }
  // the value of entry i (read after its tag was seen)
  uint64_t value(int i) const { return ((const volatile Published*)h)[i].value; }
  // Spins until [i0, i1) carry seq; a stream that finished or failed before
  // that ends the wait with an error (a finished stream's stores are all
  // visible to the host).
  void wait(int i0, int i1, uint64_t seq, hipStream_t s, const char* what) const {
    for (uint64_t spins = 1; !ready(i0, i1, seq); ++spins) {
      if (spins % 1024) continue;
      const hipError_t e = hipStreamQuery(s);
      if (e == hipErrorNotReady) continue;
      G5_HIP(e);
      G5_CHECK(ready(i0, i1, seq), what);
    }
  }
};

}  // namespace g5
