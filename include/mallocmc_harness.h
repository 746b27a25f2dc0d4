/*
 * Harness entry points of libmallocmc_hip.so: the device kernels of the
 * reference's verify_heap test (tests/verify_heap.cpp:241-405) restated for
 * offset arrays, plus the device-atomic ceiling microbenchmark that the
 * allocator's throughput is judged against.  Same conventions as
 * mallocmc_hip.h (void* stream, device pointers, 0 / negative return).
 */
#ifndef MALLOCMC_HARNESS_H
#define MALLOCMC_HARNESS_H

#include "mallocmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Writes the uint64 value (salt + i) over the first floor(d_sizes[i] / 8)
 * words of allocation i (the AllocAll fill, verify_heap.cpp:351-356). */
int mmc_write_pattern(mmc_heap* heap, const uint64_t* d_offsets, const uint32_t* d_sizes, uint64_t n, uint64_t salt,
                      void* stream);

/* Checks every allocation (Check_content, verify_heap.cpp:241-274) and the
 * pointer invariants.  d_result[0..4] (uint64, zeroed by the call):
 * pattern mismatches, misaligned offsets (not a multiple of dataAlignment),
 * allocations outside the page area, null entries with a non-zero size,
 * and the number of checked allocations. */
int mmc_verify_pattern(mmc_heap* heap, const uint64_t* d_offsets, const uint32_t* d_sizes, uint64_t n,
                       uint64_t salt, uint64_t* d_result, void* stream);

/* AllocAll (verify_heap.cpp:333-361): `blocks` x `threads` lanes call
 * malloc(size) until it returns null; each success takes the next slot of
 * d_offsets (capacity `cap`) via an atomic counter *d_count (zeroed by the
 * call).  Allocations beyond `cap` are freed again. */
int mmc_alloc_until_null(mmc_heap* heap, uint32_t size, uint64_t* d_offsets, uint64_t cap, uint64_t* d_count,
                         uint32_t blocks, uint32_t threads, void* stream);

/* Device-atomic ceiling: blocks x threads lanes each perform `iters` agent-scope
 * returning atomicAdd on uint32 words of d_buf (`words` long).  pattern 0:
 * every lane its own word (distinct lines); 1: the 64 lanes of a wave share
 * one word; 2: lanes of a wave use consecutive words of one line; 3: CAS
 * instead of add, distinct words.  *d_sink receives a checksum so the work
 * cannot be elided. */
int mmc_bench_atomics(uint32_t* d_buf, uint64_t words, uint32_t pattern, uint32_t iters, uint32_t blocks,
                      uint32_t threads, uint64_t* d_sink, void* stream);

/* Plain copy kernel of `bytes` (multiple of 16) used to calibrate the HBM
 * traffic counters. */
int mmc_bench_copy(const void* d_src, void* d_dst, uint64_t bytes, void* stream);

/* Self-test of the wave primitives the wave path is built on (DPP max, min
 * and or reductions, the DPP prefix sum), no reference counterpart:
 * `waves` wavefronts draw seeded values (from `seed`), compare each primitive
 * with a lane-by-lane loop and add their mismatches to d_result[0..3]
 * (uint32: max, min, or, prefix; zeroed by the call). */
int mmc_selftest_wave(uint32_t waves, uint64_t seed, uint32_t* d_result, void* stream);

/* Host self-test of the division by invariant divisors the heap view uses
 * (regionsize, numregions, pages per access block, pagesize), no reference
 * counterpart: `count` seeded (dividend, divisor) pairs plus edge cases
 * (powers of two and their neighbours, 0 and 2^32-1 dividends, multiples of
 * the divisor and their neighbours) against the `/` operator.  Returns the
 * number of mismatches in *mismatches.  Runs on the host only. */
int mmc_selftest_fastdiv(uint64_t seed, uint64_t count, uint64_t* mismatches);

/* Host self-test of the heap configurations and the tuning the batch kernels
 * are built for at compile time (pagesize 64 KiB or 16 KiB, regionsize 16 or
 * 32, wastefactor 2, alignment 16, default hashing; mmc_default_tuning), no
 * reference counterpart: *pinned receives how many of the four
 * configurations' heap views and the default tuning their kernels accept
 * (all five when the compile-time constants equal the ones a heap and
 * mmc_default_tuning give), *wrong how many views of other configurations
 * (another wastefactor, alignment, hashing constant, pagesize or regionsize)
 * and other tunings (one field changed) they accept (none: those run the
 * kernels that read every constant at run time).  Runs on the host only. */
int mmc_selftest_pinned_geometry(uint32_t* pinned, uint32_t* wrong);

#ifdef __cplusplus
}
#endif
#endif
