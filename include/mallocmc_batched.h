/*
 * Batched entry points of libmallocmc_hip.so: wave-cooperative malloc / free
 * (include/mallocMC/detail/wave_alloc.hpp), a persistent multi-heap step
 * kernel that serves many independent heaps in one launch, and the
 * device-resident serving loop.  They have no one-to-one counterpart in the
 * reference, whose only device entry points are per-thread
 * DeviceAllocator::malloc / free (src/include/mallocMC/device_allocator.hpp:72-88);
 * they produce heap states those entry points can keep using.
 * Conventions as in mallocmc_hip.h.
 */
#ifndef MALLOCMC_BATCHED_H
#define MALLOCMC_BATCHED_H

#include "mallocmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Tuning of the wave path. */
typedef struct mmc_tuning {
    uint32_t aggregation;  /* lanes served by one group claim: 1..64 (1 = per lane) */
    uint32_t xcd_slices;   /* 0/1: reference hash; 8: one slice of each access block per XCD */
    uint32_t persistent;   /* 0: one lane per request; 1: persistent waves pull 64-request tiles */
    uint32_t waves_per_cu; /* persistent launch size (0 = 8) */
    uint32_t class_split;  /* size classes per octave = 2^class_split (0..3) */
    uint32_t flat_single;  /* 1: claim every page on its own, two round trips each (0, the default:
                              flat pages and unformatted pages several per round trip) */
    uint32_t no_skip;      /* 1: walk every window of the region scan (0, the default: after a window
                              without candidate pages, skip ahead 256 pages per round trip) */
    uint32_t no_partition; /* 1: lane i of a non-persistent launch serves request i (0, the default:
                              each block first reorders its requests by size, its frees by page,
                              so that a wave meets few size classes / pages) */
    uint32_t block_threads; /* threads per block of the non-persistent wave malloc and free, of the
                               non-persistent multi-heap step and of the fill, the requests one
                               partition sorts: 256, 512 or 1024 (0 = 1024) */
    uint32_t fresh_single;  /* 1: claim the unformatted pages of chunk sizes with a segment hierarchy
                               one at a time, two round trips each (0, the default: several per round
                               trip; flat_single = 1 implies this) */
    uint32_t lane_requests; /* requests per lane of the non-persistent, partitioned wave malloc on a
                               heap without XMallocSIMD: 1, 2 or 4 (0 = 1).  With m > 1 a block serves
                               m x block_threads requests, a size-class group spans up to
                               m x aggregation of them, and the wave free serves two per lane */
} mmc_tuning;

int mmc_default_tuning(mmc_tuning* t);

/* Wave-cooperative counterpart of mmc_malloc_batch: same inputs and outputs,
 * same heap protocol, lanes of a wavefront grouped by size class. */
int mmc_wave_malloc_batch(mmc_heap* heap, const uint32_t* d_sizes, uint64_t* d_offsets, uint64_t n,
                          const mmc_tuning* tuning, void* stream);

/* Wave-cooperative counterpart of mmc_free_batch (frees grouped by page). */
int mmc_wave_free_batch(mmc_heap* heap, const uint64_t* d_offsets, uint64_t n, const mmc_tuning* tuning,
                        void* stream);

/* Fill until null: `lanes` lanes (a multiple of 64) each malloc in a loop
 * until the heap returns null for them -- or, with d_stop set, until the
 * first null any lane sees (it sets *d_stop, which every wave checks before
 * its next request) -- every success appended to
 * d_offsets / d_sizes at the position taken from *d_count (capacity `cap`;
 * allocations beyond it are freed again).  Request sizes are a two-part
 * uniform mixture: with probability p_small_permille/1000 uniform in
 * [small_lo, small_hi], else uniform in [large_lo, large_hi] (bytes, before
 * padding), drawn from splitmix64(seed, lane, iteration).  *d_count is not
 * reset by the call, so fills can append to an existing list.  Unless
 * tuning->no_partition, each workgroup of tuning->block_threads lanes
 * (1024 by default) reorders a round's requests by size before serving
 * them (a lane's null is reported back to it) and runs its rounds while any
 * of its lanes is active. */
typedef struct mmc_fill_desc {
    uint64_t seed;
    uint32_t lanes;
    uint32_t p_small_permille;
    uint32_t small_lo, small_hi, large_lo, large_hi;
    uint64_t cap;
    uint64_t* d_offsets;
    uint32_t* d_sizes;
    uint64_t* d_count;
    uint32_t* d_stop; /* optional: stop everyone at the first null (not reset by the call) */
} mmc_fill_desc;

int mmc_wave_fill_until_null(mmc_heap* heap, const mmc_fill_desc* desc, const mmc_tuning* tuning, void* stream);

/* ---------------------------------------------------------------- multi-heap */
typedef struct mmc_multi mmc_multi;

/* Groups `nheaps` existing heaps for one-launch service; the heaps stay owned
 * by the caller and must outlive the group. */
int mmc_multi_create(mmc_heap* const* heaps, uint32_t nheaps, mmc_multi** out);
int mmc_multi_destroy(mmc_multi* m);

/* Work description of one step over many heaps.  Slot s (0 <= s < nslots)
 * belongs to heap d_slot_heap[s]; slots of one heap are contiguous and each
 * 64-slot tile of d_tiles (uint32 first slot, tile i covers
 * [d_tiles[i], min(d_tiles[i] + 64, end of that heap's slots))) lies in one
 * heap.  An entry MMC_NO_TILE is padding: it covers no slot, so a caller can
 * start each heap's tiles at a multiple of the workgroup's block_threads/64
 * and let every workgroup of the workgroup form partition its requests.  Each step, for every slot: a live slot (d_offsets[s] != MMC_NULL)
 * frees its allocation with probability p_free_permille/1000; a dead slot
 * allocates with probability p_alloc_permille/1000 a size drawn from the
 * table d_size_table (n_sizes entries) with the cumulative weights
 * d_size_cdf (uint32, last entry 2^32-1).  Randomness is splitmix64 of
 * (seed, *d_step, slot).  *d_step is incremented by the launch, so a captured
 * graph replays successive steps.  d_counters (uint64[4], accumulated):
 * mallocs, frees, nulls, tiles processed. */
#define MMC_NO_TILE 0xffffffffu

typedef struct mmc_step_desc {
    const uint32_t* d_slot_heap;
    const uint32_t* d_tiles;
    uint32_t n_tiles;
    uint32_t n_slots;
    const uint32_t* d_size_table;
    const uint32_t* d_size_cdf;
    uint32_t n_sizes;
    uint32_t p_free_permille;
    uint32_t p_alloc_permille;
    uint32_t reserved;
    uint64_t seed;
    uint64_t* d_offsets;
    uint32_t* d_sizes;    /* size of the live allocation per slot (for verification) */
    uint64_t* d_step;
    uint64_t* d_counters;
    uint32_t* d_tile_cursor; /* uint32[2], zero before the first launch; each launch leaves it at zero */
} mmc_step_desc;

/* One step.  tuning NULL or tuning->persistent = 1: persistent waves pull the
 * tiles from d_tile_cursor[0].  tuning->persistent = 0: one workgroup of
 * tuning->block_threads threads per block_threads/64 consecutive tiles; when
 * its tiles with slots lie in one heap the workgroup first reorders its frees
 * by page and its mallocs by size (unless tuning->no_partition).  The workgroup form
 * measured faster at every fill tried; the multi-heap driver autotunes the
 * form and the workgroup size. */
int mmc_multi_step(mmc_multi* m, const mmc_step_desc* desc, const mmc_tuning* tuning, void* stream);

/* The step of mmc_multi_step over `heaps` directly, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc -- and so cannot build an mmc_multi group, whose heap
 * table mmc_multi_create allocates.  The heap table and the tile cursor live
 * in one fixed device array of the library instead (the table is written by
 * kernels whose arguments carry the heap views, so the call can be
 * captured); desc->d_tile_cursor is ignored.  Results are those of
 * mmc_multi_step with the same heaps, descriptor and tuning.
 * Limits and sharing: at most MMC_STATIC_MAX_HEAPS heaps (MMC_EINVAL beyond);
 * every call on one device uses that device's one array, so calls that
 * overlap on one device (on different streams, or from several host threads)
 * race and must be serialized by the caller.  The first call on a device
 * looks the array up and should happen outside a stream capture. */
#define MMC_STATIC_MAX_HEAPS 256u
int mmc_multi_step_static(mmc_heap* const* heaps, uint32_t nheaps, const mmc_step_desc* desc,
                          const mmc_tuning* tuning, void* stream);

#ifdef __cplusplus
}
#endif
#endif
