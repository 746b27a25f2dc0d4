/*
 * CPU oracle for the ScatterAlloc device heap -- TEST INFRASTRUCTURE ONLY.
 *
 * This is an integer-only C11 restatement of mallocMC's Scatter creation
 * policy (src/include/mallocMC/creationPolicies/Scatter.hpp) together with
 * the Noop distribution, Shrink alignment and ReturnNull OOM policies, as the
 * reference executes them on its CPU backend: warpSize = 1 and
 * laneid = warpid = smid = 0 (src/include/mallocMC/mallocMC_utils.hpp:85-87,
 * 91-137).  Each function cites the reference lines it follows.
 *
 * Only the test-suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load this library, and only as the checker or the reported
 * CPU baseline.  The product (libmallocmc_hip.so) never links or calls it.
 *
 * What pins this restatement (tests/test_oracle.py): closed forms of the
 * geometry, the page capacities and the first serialized offset of every
 * size on a fresh heap; the slot counts and first offset a probe build of the
 * reference printed for 16,384 x malloc(16) on a 64 MiB heap; brute-force streams on tiny heaps checking
 * after every step that no allocations overlap, the wastefactor rule holds, a
 * page hands out its lowest free chunk, the count equals the live chunks and a
 * null means no page can take the request (for multi-page requests: no run
 * of free pages, or _firstFreePageBased wrapped below page 0), and every page
 * above _firstFreePageBased belongs to a live multi-page run; the special
 * cases of 0 bytes, multi-page first fit from the top and the pagesize-chunk
 * leak.
 *
 * The stateful heuristics of the walk -- the region fill-level increment and
 * its cap (Scatter.hpp:612-617), the reset at P/2/c (:727-735), the
 * access-block cursor's advance, guard and atomicMin (:620-629, 733-734), the
 * 3R/4 first-pass gate and the second pass (:553, 634) and
 * _firstFreePageBased's CAS, atomicMax and uint32 wrap (:796-801, 885) -- are
 * pinned by known-answer streams whose every offset, fill level and cursor
 * value was derived by hand from Scatter.hpp (tests/oracle_kats.py).  Builds
 * with ORC_MUTANT set break one heuristic each; they are test-only, and each
 * fails its stream (tests/test_oracle_kats.py).
 *
 * Documented divergences from the reference (all cases where the reference is
 * undefined or unsafe; none is reachable by a valid workload otherwise):
 *   D1 heap < accessblocksize: the reference divides by zero
 *      (Scatter.hpp:992-993); here the number of access blocks is clamped to 1.
 *   D2 the second page-based scan reads the PTE one past the end
 *      (Scatter.hpp:835 with the inclusive loop at :785-788); here that
 *      index is treated as "not free".
 *   D3 destroy() of a pointer whose page has chunksize 0 divides by zero
 *      (Scatter.hpp:929); here such a free is ignored.
 */
#ifndef MALLOCMC_SCATTER_ORACLE_H
#define MALLOCMC_SCATTER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NULL UINT64_MAX

/* Scatter/Shrink configuration: Scatter.hpp:56-85, Shrink.hpp:60-66. */
typedef struct orc_config {
    uint32_t pagesize;
    uint32_t regionsize;
    uint64_t accessblocksize;
    uint32_t wastefactor;
    uint32_t resetfreedpages;
    uint32_t hashingK;
    uint32_t hashingDistMP;
    uint32_t hashingDistWP;
    uint32_t hashingDistWPRel;
    uint32_t dataAlignment; /* power of two; 1 == AlignmentPolicies::Noop */
    uint32_t reserved;
} orc_config;

typedef struct orc_geometry {
    uint32_t numpages;
    uint32_t numregions;
    uint32_t accessblocks;
    uint32_t pagesperblock;
    uint32_t hierarchyThreshold;
    uint32_t minChunkSize;
    uint32_t maxOnPageMasks;
    uint32_t clampedBlocks; /* 1 if divergence D1 applied */
    uint64_t ptesOffset;
    uint64_t regionsOffset;
    uint64_t usedBytes;
} orc_geometry;

/* Atomic back-ends for the concurrent CPU baseline. */
enum {
    ORC_ATOMIC_LOCKFREE = 0, /* __atomic builtins, seq_cst */
    ORC_ATOMIC_MUTEX16 = 1   /* 16-way address-hashed mutex, as alpaka AtomicStdLibLock<16> */
};

typedef struct orc_heap orc_heap;

void orc_default_config(orc_config* cfg);
/* Returns NULL and sets *err (if given) to a message on invalid input. */
orc_heap* orc_create(const orc_config* cfg, uint64_t memsize, const char** err);
void orc_destroy(orc_heap* h);
void orc_get_geometry(const orc_heap* h, orc_geometry* g);
/* Pure geometry (no allocation); returns 0 on success. */
int orc_compute_geometry(const orc_config* cfg, uint64_t memsize, orc_geometry* g);
void orc_set_atomic_mode(orc_heap* h, int mode);

/* Device-side malloc / free semantics of DeviceAllocator (device_allocator.hpp:72-88)
 * on the CPU backend.  Offsets are relative to the pool start. */
uint64_t orc_malloc(orc_heap* h, uint64_t bytes);
void orc_free(orc_heap* h, uint64_t offset);
/* Allocator::getAvailableSlots (allocator.hpp:227-233 -> Scatter.hpp:1158-1219). */
uint32_t orc_available_slots(orc_heap* h, uint64_t slotsize);

/* Serialized request stream: ops[i] >= 0 is malloc(ops[i]) and writes the
 * offset (or ORC_NULL) to out[i]; ops[i] < 0 frees the result of request
 * (-ops[i] - 1) and writes 0 to out[i]. */
void orc_run_stream(orc_heap* h, const int64_t* ops, uint64_t n, uint64_t* out);

/* Raw metadata access for tests. */
void orc_read_pte(const orc_heap* h, uint32_t page, uint32_t* chunksize, uint32_t* count, uint32_t* bitmask);
uint32_t orc_read_region(const orc_heap* h, uint32_t region);
void orc_read_scalars(const orc_heap* h, uint32_t* firstFreePageBased, uint32_t* firstfreeblock);
const uint8_t* orc_pool(const orc_heap* h);

/* Concurrent CPU baseline in the execution model of alpaka's CpuOmp2Blocks
 * backend (one thread per block, OpenMP over blocks).  Thread t mallocs
 * sizes[t], writes a uint64 pattern (t) over the buffer, verifies it, and
 * frees it.  Phase times are in seconds.  Returns the number of pattern
 * mismatches. */
typedef struct orc_bench_result {
    double t_malloc;
    double t_write_verify;
    double t_free;
    uint64_t nulls;
    uint64_t mismatches;
    int threads_used;
} orc_bench_result;
int orc_bench_concurrent(orc_heap* h, const uint32_t* sizes, uint64_t n, int nthreads, uint64_t* offsets,
                         orc_bench_result* res);

#ifdef __cplusplus
}
#endif
#endif
