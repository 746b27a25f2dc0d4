/*
 * C ABI of the MI355X-native ScatterAlloc heap (libmallocmc_hip.so).
 *
 * Plain pointers and sizes only: `stream` is a hipStream_t passed as void*
 * (NULL = default stream); every d_* pointer is device memory.  Each entry
 * point returns 0 on success and a negative MMC_E* code on failure, with a
 * message in mmc_last_error() (the reference throws std::invalid_argument from
 * its host code, src/include/mallocMC/creationPolicies/Scatter.hpp:1065-1071).
 * Device-side allocation failures are not errors: they come back as
 * MMC_NULL_OFFSET, the ReturnNull OOM policy
 * (src/include/mallocMC/oOMPolicies/ReturnNull.hpp:47-51).
 *
 * Offsets are byte offsets from the pool start returned by
 * mmc_heap_location(); C++ device code uses pointers instead (see
 * include/mallocMC/mallocMC.hpp).
 */
#ifndef MALLOCMC_HIP_H
#define MALLOCMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMC_NULL_OFFSET UINT64_MAX

#define MMC_OK 0
#define MMC_EINVAL (-1)
#define MMC_EHIP (-2)
#define MMC_ENOMEM (-3)

/* DistributionPolicies (src/include/mallocMC/distributionPolicies/) */
#define MMC_DIST_NOOP 0u
#define MMC_DIST_XMALLOCSIMD 1u

/* Flags */
#define MMC_FLAG_FIX_PAGESIZE_LEAK 1u /* release pagesize-sized chunks through the chunked path */

/* Heap configuration: ScatterConf::DefaultScatterConfig +
 * DefaultScatterHashingParams (Scatter.hpp:56-85), the XMallocSIMD choice
 * (XMallocSIMD.hpp:49-55) and Shrink's dataAlignment (Shrink.hpp:60-66;
 * 1 selects AlignmentPolicies::Noop). */
typedef struct mmc_config {
    uint32_t pagesize;
    uint32_t regionsize;
    uint64_t accessblocksize;
    uint32_t wastefactor;
    uint32_t resetfreedpages;
    uint32_t hashingK;
    uint32_t hashingDistMP;
    uint32_t hashingDistWP;
    uint32_t hashingDistWPRel;
    uint32_t dataAlignment;
    uint32_t distribution;
    uint32_t flags;
    uint32_t reserved[3];
} mmc_config;

typedef struct mmc_geometry {
    uint32_t numpages;
    uint32_t numregions;
    uint32_t accessblocks;
    uint32_t pagesperblock;
    uint32_t hierarchyThreshold;
    uint32_t minChunkSize;
    uint32_t maxOnPageMasks;
    uint32_t clampedBlocks;
    uint64_t ptesOffset;
    uint64_t regionsOffset;
    uint64_t usedBytes;
    uint32_t sideMaskWords; /* segment-mask words per page in the side array (no reference counterpart) */
    uint32_t reserved0;
} mmc_geometry;

typedef struct mmc_heap mmc_heap;

const char* mmc_last_error(void);
const char* mmc_version(void);

/* Fills the reference defaults (Scatter.hpp:56-85, Shrink.hpp:60-66, Noop distribution). */
int mmc_default_config(mmc_config* cfg);

/* Host geometry of a heap of `bytes` (Scatter.hpp:962-1014); no device work. */
int mmc_compute_geometry(const mmc_config* cfg, uint64_t bytes, mmc_geometry* out);

/* Allocator ctor (src/include/mallocMC/allocator.hpp:136-159, 179-183):
 * reserves the pool, aligns it (Shrink.hpp:109-130) and enqueues the heap
 * init kernel (Scatter.hpp:1057-1096) on `stream`.  Besides the pool the heap
 * holds a side array of segment masks (up to 128 B per page, at most 0.8% of
 * the pool for pagesize >= 16 KiB), where the reference keeps them in the last
 * bytes of each page; page geometry and offsets are the reference's.
 * Fails with MMC_EINVAL when the page table (pool + numpages * pagesize)
 * would not start 8-byte aligned: chunk size and count of a page are updated
 * as one 64-bit word (possible only for a pagesize that is not a multiple
 * of 8). */
int mmc_create(const mmc_config* cfg, uint64_t bytes, void* stream, mmc_heap** out);

/* Heap inside a pool the caller reserved (ReservePool policy,
 * reservePoolPolicies/AlpakaBuf.hpp:41-46, then AlignmentPolicy::alignPool,
 * Shrink.hpp:109-130): `pool` must already be aligned to cfg->dataAlignment,
 * and the page table must start 8-byte aligned as for mmc_create.  Enqueues the init kernel (Scatter.hpp:1057-1096) on `stream`.  The heap does
 * not own the pool: mmc_destroy leaves it to the caller. */
int mmc_create_in_pool(const mmc_config* cfg, void* pool, uint64_t bytes, void* stream, mmc_heap** out);

/* Allocator dtor (allocator.hpp:185-189).  Synchronizes the device first. */
int mmc_destroy(mmc_heap* heap);

/* Allocator::destructiveResize (allocator.hpp:195-200). */
int mmc_destructive_resize(mmc_heap* heap, uint64_t bytes, void* stream);

/* Allocator::getHeapLocations (allocator.hpp:235-241): pool start and size. */
int mmc_heap_location(const mmc_heap* heap, void** pool, uint64_t* size);

int mmc_get_config(const mmc_heap* heap, mmc_config* out);
int mmc_get_geometry(const mmc_heap* heap, mmc_geometry* out);

/* Allocator::info (allocator.hpp:214-223) and the policies' classname()
 * strings.  Writes at most `len` bytes including the terminator. */
int mmc_info(const mmc_heap* heap, char* buf, uint64_t len);

/* Allocator::getAvailableSlots (allocator.hpp:227-233 ->
 * Scatter::getAvailableSlotsHost, Scatter.hpp:1235-1295): pads the size,
 * counts on the device and blocks on `stream`. */
int mmc_available_slots(mmc_heap* heap, uint64_t slot_size, void* stream, uint32_t* out);

/* Same count left in device memory (*d_out), without host synchronization. */
int mmc_available_slots_async(mmc_heap* heap, uint64_t slot_size, uint32_t* d_out, void* stream);

/* Copies the by-value device handle (mallocMC::detail::ScatterHeap) into
 * `dst` for C++ kernels compiled against include/mallocMC; `*size` is in/out.
 * The reference's counterpart is Allocator::getAllocatorHandle
 * (allocator.hpp:202-212). */
int mmc_get_handle(const mmc_heap* heap, void* dst, uint64_t* size);

/* One launch in which lane i performs the device-side
 * DeviceAllocator::malloc(acc, d_sizes[i]) (device_allocator.hpp:72-82) and
 * stores the pool offset in d_offsets[i] (MMC_NULL_OFFSET on OOM or 0 bytes). */
int mmc_malloc_batch(mmc_heap* heap, const uint32_t* d_sizes, uint64_t* d_offsets, uint64_t n, void* stream);

/* One launch in which lane i performs DeviceAllocator::free(acc, p)
 * (device_allocator.hpp:84-88) of the pointer at d_offsets[i];
 * MMC_NULL_OFFSET entries are skipped, as free(nullptr) is. */
int mmc_free_batch(mmc_heap* heap, const uint64_t* d_offsets, uint64_t n, void* stream);

/* Serialized request stream on one lane with the reference CPU backend's ids
 * (warpSize 1, laneid = warpid = smid = 0; mallocMC_utils.hpp:85-87):
 * d_ops[i] >= 0 is malloc(d_ops[i]) with the offset in d_out[i];
 * d_ops[i] < 0 frees the result of request (-d_ops[i] - 1), d_out[i] = 0. */
int mmc_serial_stream(mmc_heap* heap, const int64_t* d_ops, uint64_t* d_out, uint64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
