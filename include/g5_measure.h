/*
 * g5_measure.h -- measurement helpers of libg5.so (no counterpart in the
 * reference; they carry none of its arithmetic).  They exist to calibrate the
 * hardware counters behind the roofline figures: a profiler's memory-traffic
 * counters are read once for a kernel whose traffic is known exactly, and the
 * ratio scales what the same counters read for the BFS kernels.
 */
#ifndef G5_MEASURE_H
#define G5_MEASURE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random 8-byte reads: allocates array_bytes of device memory on `device`
 * (a multiple of 8, at least 8 bytes), fills it, then times one launch of
 * k_measure_gather, which reads `gathers` 8-byte words at pseudo-random
 * positions of the array (a 64-bit mix of the read index, uniform over the
 * array; no other memory traffic except one 8-byte store per workgroup).
 * *ms receives the launch's device time (HIP events), *checksum the XOR of
 * every word read (so no read can be dropped).  Frees everything before it
 * returns.  0 on success, a negative G5_ERR_* code otherwise. */
int g5_measure_gather(int device, int64_t array_bytes, int64_t gathers, double* ms, uint64_t* checksum);

#ifdef __cplusplus
}
#endif

#endif
