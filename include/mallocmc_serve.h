/*
 * Device-resident serving loop over the heap (libmallocmc_hip.so).  No
 * counterpart in the reference, which has no request loop; it exercises
 * device-side malloc/free (src/include/mallocMC/device_allocator.hpp:72-88)
 * the way a continuous-batching server does.
 *
 * Each request holds a growing paged buffer: `block_bytes` blocks, each holding
 * `tokens_per_block` tokens of `token_bytes`.  One step:
 *   1. admission: requests whose arrival step has come are admitted in arrival
 *      order while a request slot is free and the reserved bytes (projected
 *      blocks of every admitted request) stay within `budget_bytes`;
 *   2. decode: every active request appends one token, taking a new block from
 *      the heap when its last block is full (wave-cooperative malloc), and
 *      writes the token's payload words;
 *   3. completion: a request that reached its length reads its whole buffer
 *      back into a 64-bit checksum and frees its blocks (wave-cooperative
 *      free); its slot and reservation came back in decode.
 * A step is two launches: decode, which also returns the slots and
 * reservations of the requests it finishes, then completion, whose first
 * workgroup runs the admission for the next step (the first admission runs
 * at load).  Both can be captured in one graph; the step counter lives in
 * device memory.  Payload word w of token t of request r is
 * mmc_serve_word(r, t, w) below, so every checksum has a closed form.
 */
#ifndef MALLOCMC_SERVE_H
#define MALLOCMC_SERVE_H

#include "mallocmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mmc_serve mmc_serve;

typedef struct mmc_serve_config {
    uint32_t max_slots;        /* concurrent requests (multiple of 64) */
    uint32_t block_bytes;      /* bytes per paged-buffer block (e.g. 4096) */
    uint32_t tokens_per_block; /* tokens per block (e.g. 16) */
    uint32_t token_bytes;      /* bytes per token, multiple of 16 (tokens are written as 16-byte words; e.g. 256) */
    uint32_t max_tokens;       /* longest request (e.g. 8192) */
    uint32_t reserved;
    uint64_t budget_bytes;     /* admission budget in heap bytes */
} mmc_serve_config;

typedef struct mmc_serve_stats {
    uint64_t step;             /* steps run */
    uint64_t admitted;
    uint64_t completed;
    uint64_t block_mallocs;
    uint64_t block_frees;
    uint64_t nulls;            /* block mallocs that returned null */
    uint64_t queue_wait_steps; /* sum over admitted requests of admit - arrival */
    uint64_t active;           /* requests active after the last step */
    uint64_t reserved_bytes;
    uint64_t max_active;
} mmc_serve_stats;

/* Payload word w of token t of request r (integer hash, identical on host). */
static inline uint32_t mmc_serve_word(uint32_t r, uint32_t t, uint32_t w)
{
    uint32_t x = r * 0x9E3779B1u ^ (t * 0x85EBCA77u) ^ (w * 0xC2B2AE3Du);
    x ^= x >> 15;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    return x;
}

/* Fails with MMC_EINVAL unless max_slots is a multiple of 64, token_bytes a
 * multiple of 16, tokens_per_block * token_bytes <= block_bytes, and
 * 64 * ceil(max_tokens / tokens_per_block) * block_bytes < 2^32 (the
 * admission adds up 64 requests' projected bytes in 32 bits). */
int mmc_serve_create(mmc_heap* heap, const mmc_serve_config* cfg, mmc_serve** out);
int mmc_serve_destroy(mmc_serve* s);

/* Loads a trace of n requests sorted by arrival step (device arrays):
 * d_arrival[i] = arrival step, d_length[i] = tokens (1..max_tokens),
 * d_ids[i] = the request id its payload is derived from.
 * d_checksums (n entries) receives each request's checksum at completion:
 * sum over k of word(id, k / W, k % W) * (2k + 1) mod 2^64, W = token_bytes/4,
 * k = 0 .. length*W - 1, read back from the paged buffer.
 * d_admit_step (n entries, optional) receives its admission step.  Resets the
 * loop (request slots, counters, step 0); the heap must hold no blocks of an
 * earlier trace.  Checks the trace on the device and blocks on `stream` for
 * the result: a length outside 1..max_tokens or an arrival step below its
 * predecessor's fails with MMC_EINVAL naming the first such request, and
 * nothing is admitted. */
int mmc_serve_load(mmc_serve* s, const uint32_t* d_arrival, const uint32_t* d_length, const uint32_t* d_ids,
                   uint32_t n, uint64_t* d_checksums, uint32_t* d_admit_step, void* stream);

/* Enqueues one step (decode, then completion with the next admission) on
 * `stream`.  With a timestamp buffer attached, the decode kernel's first
 * workgroup records the device clock into the next entry of that buffer
 * before any other work of the step. */
int mmc_serve_step(mmc_serve* s, void* stream);

/* Attaches a device buffer of `capacity` uint64 device-clock stamps (NULL
 * detaches): every later step, and each mmc_serve_stamp call, appends one
 * stamp at its start, so step i took stamps[i + 1] - stamps[i] clock ticks
 * on the device, with no host launch in between when the steps are replayed
 * from one graph.  Stamps beyond `capacity` are dropped; the count restarts
 * at mmc_serve_load.  *clock_hz receives the clock's frequency. */
int mmc_serve_timestamps(mmc_serve* s, uint64_t* d_stamps, uint32_t capacity, uint64_t* clock_hz);

/* Appends one stamp (e.g. after the last step) on `stream`. */
int mmc_serve_stamp(mmc_serve* s, void* stream);

/* Number of stamps recorded since mmc_serve_load (synchronizes `stream`). */
int mmc_serve_stamp_count(mmc_serve* s, uint32_t* count, void* stream);

/* Copies the counters to the host (synchronizes `stream`). */
int mmc_serve_get_stats(mmc_serve* s, mmc_serve_stats* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
