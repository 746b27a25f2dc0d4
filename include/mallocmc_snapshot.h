/*
 * Heap snapshots and their offline verification (libmallocmc_hip.so).
 *
 * No reference counterpart: the reference checks a heap only from inside its
 * own kernels (tests/verify_heap.cpp:241-405).  A snapshot copies a quiescent
 * heap's metadata -- PTEs, region fill levels, segment masks and the three
 * heap scalars, optionally the page payload -- to host memory in a
 * relocatable format: everything is page indices and pool offsets, no device
 * addresses, so a snapshot can be saved, moved and checked on any host.
 *
 * Layout (little endian, sections 8-byte aligned, offsets from the start):
 *   mmc_snapshot_header | PTEs (3 x uint32 per page: chunksize, count,
 *   bitmask) | region fill levels (uint32 per region) | segment masks
 *   (geom.sideMaskWords x uint32 per page) | payload (pagesize bytes per
 *   page, only with MMC_SNAPSHOT_PAYLOAD).
 *
 * Conventions as in mallocmc_hip.h.
 */
#ifndef MALLOCMC_SNAPSHOT_H
#define MALLOCMC_SNAPSHOT_H

#include "mallocmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MMC_SNAPSHOT_MAGIC "MMCSNAP1"
#define MMC_SNAPSHOT_PAYLOAD 1u /* include the page payload */

typedef struct mmc_snapshot_header {
    char magic[8];       /* MMC_SNAPSHOT_MAGIC, not NUL terminated */
    uint32_t version;    /* 1 */
    uint32_t flags;      /* MMC_SNAPSHOT_* */
    mmc_config cfg;
    mmc_geometry geom;
    uint64_t memsize;    /* usable pool bytes after alignment */
    uint32_t firstFreePageBased, firstfreeblock, pagebasedMutex, reserved;
    uint64_t off_ptes, off_regions, off_masks, off_payload; /* off_payload 0 without payload */
    uint64_t total_bytes;
} mmc_snapshot_header;

/* Bytes a snapshot of `heap` with `flags` takes. */
int mmc_snapshot_bytes(const mmc_heap* heap, uint32_t flags, uint64_t* bytes);

/* Copies the heap's state into `dst` (host memory, `capacity` bytes) after
 * the work already enqueued on `stream`; blocks until the copy is done, as
 * Allocator::getAvailableSlots blocks on its queue (allocator.hpp:227-233).
 * The heap must be quiescent: no kernel may allocate or free meanwhile. */
int mmc_snapshot(mmc_heap* heap, void* dst, uint64_t capacity, uint32_t flags, void* stream);

/* Findings of mmc_snapshot_verify.  Error codes, by first occurrence:
 *   1 free page with a non-zero count, bits or segment masks
 *   2 chunk bits or masks outside the page's capacity, or masks on a flat page
 *   3 count differs from the number of chunks marked in use
 *   4 broken multi-page run (a page of the run has another chunk size)
 *   5 heap scalars out of range
 *   6 live allocation outside the page area, on a free page or misaligned
 *   7 live allocation whose chunk bit is clear, or larger than its chunk
 *   8 two live allocations on one chunk / run
 *   9 (complete live list) a page's count differs from its live chunks,
 *     or an allocated run has no live allocation
 *  10 malformed snapshot (magic, version, sizes)
 * Warnings do not count as errors: segment hints set on a segment that is
 * not full (a benign race of the reference protocol, Scatter.hpp:408-419),
 * and, with resetfreedpages, chunked pages left formatted with count 0. */
typedef struct mmc_verify_report {
    uint64_t pages;            /* pages checked */
    uint64_t free_pages;       /* chunk size 0 */
    uint64_t chunked_pages;    /* 0 < chunk size <= pagesize */
    uint64_t run_pages;        /* pages of multi-page runs */
    uint64_t runs;
    uint64_t used_chunks;      /* chunks marked in use */
    uint64_t empty_formatted;  /* warning: chunked pages with count 0 */
    uint64_t stale_hints;      /* warning: hint bits on segments that are not full */
    uint64_t live_checked;
    uint64_t errors;
    uint64_t first_error_page; /* page index, or live index for codes 6-9 */
    uint32_t first_error_code;
    uint32_t reserved;
    char message[160];         /* description of the first error */
} mmc_verify_report;

#define MMC_VERIFY_LIVE_COMPLETE 1u /* the live list holds every allocation of the heap */
#define MMC_VERIFY_LIVE_INTERIOR 2u /* live offsets may point inside a chunk (XMallocSIMD slices) */

/* Checks a snapshot on the host: per-page metadata invariants, multi-page
 * runs, heap scalars, and, given live allocations (host arrays of pool
 * offsets and byte sizes, n_live entries, may be NULL / 0), that each one is
 * marked allocated, fits its chunk and is not shared.  Returns MMC_OK when
 * the snapshot could be read (the report says whether it is consistent). */
int mmc_snapshot_verify(const void* snapshot, uint64_t bytes, const uint64_t* live_offsets,
                        const uint32_t* live_sizes, uint64_t n_live, uint32_t flags, mmc_verify_report* report);

#ifdef __cplusplus
}
#endif
#endif
