// Device-resident serving loop: admission, decode (growing paged buffers on
// the heap) and completion kernels.  Declarations: include/mallocmc_serve.h.
#include "mallocmc_serve.h"

#include "mallocMC/detail/wave_alloc.hpp"
#include "mmc_internal.hpp"

#include <string>

using namespace mallocMC;
using namespace mallocMC::detail;

namespace
{
    constexpr u32 kEmpty = 0xffffffffu;

    enum Counter : u32
    {
        cAdmitted = 0,
        cCompleted,
        cMallocs,
        cFrees,
        cNulls,
        cQueueWait,
        cActive,
        cMaxActive,
        cCount
    };

    struct ServeScalars
    {
        u64 step;     // admission rounds done (the next admission serves arrivals <= step)
        u64 reserved;
        u32 cursor;
        u32 freeTop;
        u32 nDone[2];   // entries of the two done lists (a step uses list step & 1)
        u32 nChunks[2]; // checksum chunks of their requests (decode adds them)
        u64 counters[cCount];
        u32 nStamps;  // device-clock stamps recorded since load
        u32 parity;   // the done list this step's decode filled (completion reads it)
        u32 firstBad; // index of the first request a load rejected (UINT32_MAX: none)
        u32 pad;
    };

    constexpr u32 kTokensPerChunk = 64u;

    //! Everything a serving kernel needs, passed by value.
    struct ServeView
    {
        ScatterHeap heap;
        ServeScalars* sc;
        u32* freeStack;
        u32* slotReq;
        u32* slotLen;
        u32* slotCur;
        u64* slotBlocks;
        u32* doneList; // slots of the requests finished in a step, in decode's order
        u32* doneLen;  // their lengths and request indices, so that completion
        u32* doneReq;  // reads an entry in one round trip; two lists of maxSlots
                       // entries each, alternating by step
        const u32* arrival;
        const u32* length;
        const u32* ids;
        u64* checksums;
        u32* admitStep;
        u32 nRequests;
        u32 maxSlots;
        u32 blockBytes;
        u32 tokensPerBlock;
        u32 tokenBytes;
        u32 maxBlocks;
        u64 budget;
        u64* stamps; // device-clock stamp per step (null: none)
        u32 stampCapacity;
    };

    __device__ __forceinline__ u32 payloadWord(u32 r, u32 t, u32 w)
    {
        u32 x = r * 0x9E3779B1u ^ (t * 0x85EBCA77u) ^ (w * 0xC2B2AE3Du);
        x ^= x >> 15;
        x *= 0x2C1B3C6Du;
        x ^= x >> 12;
        return x;
    }

    //! One wave: admit the requests that have arrived by step sc->step, in
    //! arrival order, while a slot and budget are available; then advance
    //! the step.  Runs at the start of the previous step's completion kernel
    //! (or at load for step 0): that step's decode already returned the
    //! slots and reservations of the requests it finished.  Up to 64
    //! candidates per round: their projected bytes are summed with a prefix
    //! scan, the leading run that fits the budget and the free slots is
    //! admitted at once, so a round costs a few round trips however many
    //! requests it admits.  Slot state is written with device-scope atomics.
    __device__ void admitWave(const ServeView& v)
    {
        ServeScalars* sc = v.sc;
        const u32 lane = hw::laneid();
        u64 step = 0u, reserved = 0u;
        u32 cursor = 0u, top = 0u;
        if(lane == 0u)
        {
            step = atom::load(&sc->step);
            cursor = atom::load(&sc->cursor);
            reserved = atom::load(&sc->reserved);
            top = atom::load(&sc->freeTop);
        }
        step = hw::readlane64(step, 0u);
        reserved = hw::readlane64(reserved, 0u);
        cursor = wave::lane(cursor, 0u);
        top = wave::lane(top, 0u);
        u64 admitted = 0u, wait = 0u;
        for(;;)
        {
            const u32 idx = cursor + lane;
            const bool arrived = idx < v.nRequests && v.arrival[idx] <= step;
            const u32 len = arrived ? v.length[idx] : 0u;
            const u32 need = arrived ? divup(len, v.tokensPerBlock) * v.blockBytes : 0u;
            // inclusive prefix of the projected bytes over lanes 0..lane
            const u32 incl = wave::inclusiveSum(need);
            const bool fits = arrived && reserved + incl <= v.budget && lane < top;
            // the leading run of lanes that fit is admitted, in order
            const u64 ok = hw::ballot(fits);
            const u32 count = ~ok == 0u ? 64u : static_cast<u32>(__builtin_ctzll(~ok));
            if(lane < count)
            {
                const u32 slot = atom::load(&v.freeStack[top - 1u - lane]);
                atom::store(&v.slotLen[slot], len);
                atom::store(&v.slotCur[slot], 0u);
                atom::store(&v.slotReq[slot], idx);
                atom::exch(&v.checksums[idx], u64{0});
                if(v.admitStep)
                    v.admitStep[idx] = static_cast<u32>(step);
            }
            const u32 waitLane = lane < count ? static_cast<u32>(step - v.arrival[idx]) : 0u;
            const u32 w = wave::sum(waitLane);
            const u32 bytes = count ? wave::lane(incl, count - 1u) : 0u;
            reserved += bytes;
            top -= count;
            cursor += count;
            admitted += count;
            wait += w;
            if(count < 64u)
                break;
        }
        if(lane == 0u)
        {
            atom::store(&sc->cursor, cursor);
            atom::exch(&sc->reserved, reserved);
            atom::store(&sc->freeTop, top);
            atom::add(&sc->counters[cAdmitted], admitted);
            atom::add(&sc->counters[cQueueWait], wait);
            const u64 active = v.maxSlots - top;
            atom::exch(&sc->counters[cActive], active);
            atom::max(&sc->counters[cMaxActive], active);
            atom::exch(&sc->step, step + 1u);
        }
    }

    //! Step part 1, one lane per request slot: append one token.
    //! Appends the device clock to the stamp buffer (one lane).
    __device__ void stamp(ServeScalars* sc, u64* stamps, u32 capacity)
    {
        const u64 t = wall_clock64();
        const u32 i = atom::add(&sc->nStamps, 1u);
        if(i < capacity)
            stamps[i] = t;
    }

    //! One wave per workgroup: the step's few hundred decode waves then land
    //! on as many CUs (with four waves per workgroup they filled only a
    //! quarter of the 256 CUs, and the step took 61 us at the median instead
    //! of 47 us).
    constexpr u32 kDecodeThreads = 64u;

    //! Geo: the heap configuration the kernel is built for (PinnedGeometry,
    //! or RuntimeGeometry for any).
    template<class Geo>
    __global__ __launch_bounds__(kDecodeThreads) void decodeKernel(ServeView view)
    {
        ServeView v = view;
        v.heap = Geo::apply(view.heap);
        // the step's start stamp: the first workgroup, before any work
        if(v.stamps && blockIdx.x == 0u && threadIdx.x == 0u)
            stamp(v.sc, v.stamps, v.stampCapacity);
        const u32 slot = blockIdx.x * blockDim.x + threadIdx.x;
        const bool valid = slot < v.maxSlots;
        // the step (its parity picks the done list), read beside the slot's
        // state in one round trip (a free slot's position and length are not
        // used), then the request id and the current block in the next,
        // issued before the malloc's own loads
        u64 stepNow = 0u;
        if(hw::laneid() == 0u)
            stepNow = atom::load(&v.sc->step);
        u32 req = kEmpty, t = 0u, len = 0u;
        if(valid)
        {
            req = v.slotReq[slot];
            t = v.slotCur[slot];
            len = v.slotLen[slot];
        }
        const bool active = req != kEmpty;
        const bool needBlock = active && (t % v.tokensPerBlock) == 0u;
        u64* blocks = v.slotBlocks + static_cast<u64>(slot) * v.maxBlocks;
        u32 id = 0u;
        u64 off = kNullOffset;
        if(active)
        {
            id = v.ids[req];
            if(!needBlock)
                off = blocks[t / v.tokensPerBlock];
        }
        const u32 par = static_cast<u32>(hw::readlane64(stepNow, 0u)) & 1u;
        if(blockIdx.x == 0u && threadIdx.x == 0u)
        {
            // the other list was read by the previous step's completion and
            // is filled by the next step's decode
            atom::store(&v.sc->nDone[par ^ 1u], 0u);
            atom::store(&v.sc->nChunks[par ^ 1u], 0u);
            atom::store(&v.sc->parity, par);
        }
        // a request that reaches its length this step takes its done-list
        // entry and returns its slot and reservation now: the returning
        // atomics travel with the malloc's loads instead of adding round
        // trips after them.  The slot may be handed out again by this step's
        // admission (in completion), which writes the slot's request,
        // length and position but not its block list; completion reads the
        // finished request from the done list and its blocks from the block
        // list, and frees the blocks.
        const bool finishing = active && t + 1u == len;
        u32 doneAt = 0u, freeAt = 0u;
        const u32 list = par * v.maxSlots;
        if(finishing)
        {
            doneAt = atom::add(&v.sc->nDone[par], 1u);
            atom::add(&v.sc->nChunks[par], divup(len, kTokensPerChunk));
            freeAt = atom::add(&v.sc->freeTop, 1u);
            atom::sub(&v.sc->reserved, static_cast<u64>(divup(len, v.tokensPerBlock)) * v.blockBytes);
            atom::add(&v.sc->counters[cCompleted], u64{1});
        }
        const WaveAlloc wa{v.heap, WaveTuning{}};
        void* p = wa.malloc<NoopDistribution>(needBlock ? v.blockBytes : 0u);
        u32 nulls = 0u;
        if(needBlock)
        {
            off = v.heap.offsetOf(p);
            blocks[t / v.tokensPerBlock] = off;
            if(p == nullptr)
                nulls = 1u;
        }
        if(active)
        {
            if(off != kNullOffset)
            {
                u32* dst = reinterpret_cast<u32*>(v.heap.pages + off
                                                  + static_cast<u64>(t % v.tokensPerBlock) * v.tokenBytes);
                const u32 words = v.tokenBytes / 4u;
                for(u32 w = 0; w < words; w += 4u)
                {
                    uint4 q;
                    q.x = payloadWord(id, t, w);
                    q.y = payloadWord(id, t, w + 1u);
                    q.z = payloadWord(id, t, w + 2u);
                    q.w = payloadWord(id, t, w + 3u);
                    *reinterpret_cast<uint4*>(dst + w) = q;
                }
            }
            v.slotCur[slot] = t + 1u;
            if(finishing)
            {
                v.doneList[list + doneAt] = slot;
                v.doneLen[list + doneAt] = len;
                v.doneReq[list + doneAt] = req;
                atom::store(&v.slotReq[slot], kEmpty);
                atom::store(&v.freeStack[freeAt], slot);
            }
        }
        const u32 mallocs = wave::sum(needBlock ? 1u : 0u);
        nulls = wave::sum(nulls);
        if(hw::laneid() == 0u)
        {
            if(mallocs)
                atom::add(&v.sc->counters[cMallocs], static_cast<u64>(mallocs));
            if(nulls)
                atom::add(&v.sc->counters[cNulls], static_cast<u64>(nulls));
        }
    }

    //! Checksum workgroups per step, beside the admission workgroup (a step
    //! finishes a few hundred 64-token chunks; one workgroup takes every
    //! 512th chunk): 512 measured p50 31.1 us, p99 45.8 us, against 32.3 /
    //! 48.0 us with 1024 and 31.0 / 46.4 us with 256.
    constexpr u32 kCompleteBlocks = 512u;

    //! Frees a finished request's blocks (wave-cooperative free); the whole
    //! wave calls it.  Nothing in the completion launch allocates, so the
    //! frees need not come before the admission: the next step's decode
    //! (the next launch) is the first to reuse the blocks.
    __device__ void freeBlocks(const ServeView& v, u32 slot, u32 len)
    {
        const u32 lane = hw::laneid();
        const WaveAlloc wa{v.heap, WaveTuning{}};
        const u32 nb = divup(len, v.tokensPerBlock);
        const u64* blocks = v.slotBlocks + static_cast<u64>(slot) * v.maxBlocks;
        u32 freed = 0u;
        for(u32 b = 0; b < nb; b += 64u)
        {
            const u32 j = b + lane;
            const u64 off = j < nb ? blocks[j] : kNullOffset;
            wa.free(off == kNullOffset ? nullptr : v.heap.pages + off);
            freed += off != kNullOffset ? 1u : 0u;
        }
        freed = wave::sum(freed);
        if(lane == 0u)
            atom::add(&v.sc->counters[cFrees], static_cast<u64>(freed));
    }

    //! Step part 2, completion: the paged buffers of the finished requests,
    //! cut into 64-token chunks spread over the whole grid, are read back into
    //! their checksums (lane = word of a token, the 4 waves of a workgroup take
    //! every fourth token, all 16 of a wave's tokens in flight at once;
    //! partial sums add atomically).  Decode already returned the finished
    //! requests' slots and reservations, so workgroup 0 admits the next
    //! step's arrivals at once (and reads no chunk), and one step is two
    //! launches.  The workgroup
    //! holding a request's first chunk frees the request's blocks beside the
    //! reads.  Workgroups without a chunk (decode added up the step's chunk
    //! count) leave at once and touch no shared word: a ticket taken by every
    //! workgroup of the grid, on one address, cost about 7 us of the step.
    template<class Geo>
    __global__ __launch_bounds__(256) void completeKernel(ServeView view)
    {
        ServeView v = view;
        v.heap = Geo::apply(view.heap);
        __shared__ u64 partial[4];
        const u32 lane = hw::laneid();
        const u32 q = threadIdx.x / 64u;
        // decode returned every finished request's slot and reservation, so
        // the first wave admits the next step's arrivals at once
        // workgroup 0 admits and does nothing else, so no chunk waits for it
        if(blockIdx.x == 0u)
        {
            if(q == 0u)
                admitWave(v);
            return;
        }
        const u32 first = blockIdx.x - 1u, nWork = gridDim.x - 1u;
        // the done list decode filled (both lists' counts in the same round
        // trip as the parity; admission does not touch them)
        const u32 par = atom::load(&v.sc->parity);
        const u32 n0 = atom::load(&v.sc->nDone[0]), n1 = atom::load(&v.sc->nDone[1]);
        const u32 c0 = atom::load(&v.sc->nChunks[0]), c1 = atom::load(&v.sc->nChunks[1]);
        const u32 nDone = par != 0u ? n1 : n0;
        const u32 total = par != 0u ? c1 : c0;
        const u32 list = par * v.maxSlots;
        if(first >= total)
            return; // no chunk here
        const u32 wpt = v.tokenBytes / 4u;
        u32 i = 0u, base = 0u; // done-list entry and its first chunk id
        for(u32 c = first; c < total; c += nWork)
        {
            // the entry holding chunk c: 64 entries per round trip (lane k
            // reads entry i + k: slot, length and request side by side), a
            // prefix sum of their chunk counts and a ballot, instead of one
            // dependent load pair per entry
            u32 nChunks = 0u, slot = 0u, len = 0u, req = 0u;
            bool found = false;
            while(i < nDone)
            {
                const u32 e = i + lane;
                const u32 es = e < nDone ? v.doneList[list + e] : 0u;
                const u32 el = e < nDone ? v.doneLen[list + e] : 0u;
                const u32 er = e < nDone ? v.doneReq[list + e] : 0u;
                const u32 n = divup(el, kTokensPerChunk);
                const u32 incl = wave::inclusiveSum(n);
                const u64 past = hw::ballot(c < base + incl); // entries whose range ends after c
                if(past != 0u)
                {
                    const u32 k = hw::ffs(past) - 1u;
                    nChunks = wave::lane(n, k);
                    slot = wave::lane(es, k);
                    len = wave::lane(el, k);
                    req = wave::lane(er, k);
                    base += wave::lane(incl, k) - nChunks;
                    i += k;
                    found = true;
                    break;
                }
                base += wave::lane(incl, 63u);
                i += 64u;
            }
            if(!found)
                break;
            const u64* blocks = v.slotBlocks + static_cast<u64>(slot) * v.maxBlocks;
            if(c == base && q == 3u)
            {
                // the request's first chunk: its last wave frees the
                // request's blocks right away, beside the checksum reads
                // (nothing in this launch allocates, and a free leaves the
                // payload as it is, so the reads stay valid)
                freeBlocks(v, slot, len);
            }
            const u32 t0 = (c - base) * kTokensPerChunk;
            const u32 t1 = umin32(len, t0 + kTokensPerChunk);
            // wave q reads tokens t0 + q, t0 + q + 4, ...: all block offsets
            // of its tokens in one round trip, then all payload words in the next
            constexpr u32 kPerWave = kTokensPerChunk / 4u;
            u64 offs[kPerWave];
#pragma unroll
            for(u32 j = 0; j < kPerWave; ++j)
            {
                const u32 tt = t0 + q + 4u * j;
                offs[j] = tt < t1 ? blocks[tt / v.tokensPerBlock] : kNullOffset;
            }
            u64 cs = 0u;
            for(u32 w = lane; w < wpt; w += 64u)
            {
                u32 x[kPerWave];
#pragma unroll
                for(u32 j = 0; j < kPerWave; ++j)
                {
                    const u32 tt = t0 + q + 4u * j;
                    x[j] = offs[j] == kNullOffset
                        ? 0u
                        : *reinterpret_cast<const u32*>(
                            v.heap.pages + offs[j] + static_cast<u64>(tt % v.tokensPerBlock) * v.tokenBytes + 4u * w);
                }
#pragma unroll
                for(u32 j = 0; j < kPerWave; ++j)
                    cs += static_cast<u64>(x[j]) * (2u * (static_cast<u64>(t0 + q + 4u * j) * wpt + w) + 1u);
            }
            for(int d = 32; d > 0; d >>= 1)
                cs += __shfl_xor(cs, d, 64);
            if(lane == 0u)
                partial[q] = cs;
            __syncthreads();
            if(threadIdx.x == 0u)
                atom::add(&v.checksums[req], partial[0] + partial[1] + partial[2] + partial[3]);
            __syncthreads();
        }
    }

    __global__ void resetKernel(ServeView v)
    {
        const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
        const u32 stride = gridDim.x * blockDim.x;
        for(u32 i = tid; i < v.maxSlots; i += stride)
        {
            v.freeStack[i] = v.maxSlots - 1u - i;
            v.slotReq[i] = kEmpty;
        }
        if(tid == 0u)
        {
            v.sc->step = 0u;
            v.sc->reserved = 0u;
            v.sc->cursor = 0u;
            v.sc->freeTop = v.maxSlots;
            v.sc->nDone[0] = 0u;
            v.sc->nDone[1] = 0u;
            v.sc->nChunks[0] = 0u;
            v.sc->nChunks[1] = 0u;
            v.sc->parity = 0u;
            v.sc->nStamps = 0u;
            v.sc->firstBad = 0xffffffffu;
            for(u32 c = 0; c < cCount; ++c)
                v.sc->counters[c] = 0u;
        }
    }

    //! Finds the first request with a length outside 1..maxTokens or an
    //! arrival step below its predecessor's (the loop's preconditions).
    __global__ void validateKernel(ServeView v, u32 maxTokens)
    {
        const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
        const u32 stride = gridDim.x * blockDim.x;
        for(u32 i = tid; i < v.nRequests; i += stride)
        {
            const u32 len = v.length[i];
            const bool bad = len == 0u || len > maxTokens || (i > 0u && v.arrival[i] < v.arrival[i - 1u]);
            if(bad)
                atom::min(&v.sc->firstBad, i);
        }
    }

    //! A stamp outside a step (e.g. after the last one).
    __global__ void stampKernel(ServeScalars* sc, u64* stamps, u32 capacity)
    {
        if(blockIdx.x == 0u && threadIdx.x == 0u)
            stamp(sc, stamps, capacity);
    }

    //! Admission for step 0 after the reset (one wave).
    __global__ void firstAdmitKernel(ServeView v)
    {
        admitWave(v);
    }
} // namespace

struct mmc_serve
{
    mmc_heap* heap = nullptr;
    mmc_serve_config cfg{};
    ServeView view{};
    void* arena = nullptr;
    u64* stamps = nullptr;
    u32 stampCapacity = 0;
};

extern "C" {

int mmc_serve_create(mmc_heap* heap, const mmc_serve_config* cfg, mmc_serve** out)
{
    if(!heap || !cfg || !out)
        return mmc::fail(MMC_EINVAL, "null argument");
    *out = nullptr;
    if(cfg->max_slots == 0 || cfg->max_slots % 64u != 0 || cfg->block_bytes == 0 || cfg->tokens_per_block == 0
       || cfg->token_bytes == 0 || cfg->token_bytes % 16u != 0 || cfg->max_tokens == 0
       || static_cast<u64>(cfg->tokens_per_block) * cfg->token_bytes > cfg->block_bytes)
        return mmc::fail(MMC_EINVAL, "bad serving configuration");
    const u32 maxBlocks = divup(cfg->max_tokens, cfg->tokens_per_block);
    if(64u * static_cast<u64>(maxBlocks) * cfg->block_bytes > 0xffffffffull)
        return mmc::fail(MMC_EINVAL, "64 x ceil(max_tokens / tokens_per_block) x block_bytes must stay below 2^32: "
                                     "admission adds 64 requests' projected bytes in 32 bits");
    mmc_serve* s = new mmc_serve();
    s->heap = heap;
    s->cfg = *cfg;
    const u64 n = cfg->max_slots;
    const u64 bytes = sizeof(ServeScalars) + 4u * n * 10u + 8u * n * maxBlocks + 256u;
    hipError_t err = hipMalloc(&s->arena, bytes);
    if(err != hipSuccess)
    {
        delete s;
        return mmc::hipFail(err, "hipMalloc(serving state)");
    }
    char* a = static_cast<char*>(s->arena);
    ServeView& v = s->view;
    v.heap = heap->view;
    v.sc = reinterpret_cast<ServeScalars*>(a);
    a += (sizeof(ServeScalars) + 255u) & ~u64{255u};
    v.slotBlocks = reinterpret_cast<u64*>(a);
    a += 8u * n * maxBlocks;
    v.freeStack = reinterpret_cast<u32*>(a);
    a += 4u * n;
    v.slotReq = reinterpret_cast<u32*>(a);
    a += 4u * n;
    v.slotLen = reinterpret_cast<u32*>(a);
    a += 4u * n;
    v.slotCur = reinterpret_cast<u32*>(a);
    a += 4u * n;
    v.doneList = reinterpret_cast<u32*>(a);
    a += 8u * n;
    v.doneLen = reinterpret_cast<u32*>(a);
    a += 8u * n;
    v.doneReq = reinterpret_cast<u32*>(a);
    v.maxSlots = cfg->max_slots;
    v.blockBytes = cfg->block_bytes;
    v.tokensPerBlock = cfg->tokens_per_block;
    v.tokenBytes = cfg->token_bytes;
    v.maxBlocks = maxBlocks;
    v.budget = cfg->budget_bytes;
    *out = s;
    return MMC_OK;
}

int mmc_serve_destroy(mmc_serve* s)
{
    if(!s)
        return MMC_OK;
    (void) hipDeviceSynchronize();
    if(s->arena)
        (void) hipFree(s->arena);
    delete s;
    return MMC_OK;
}

int mmc_serve_load(mmc_serve* s, const uint32_t* d_arrival, const uint32_t* d_length, const uint32_t* d_ids,
                   uint32_t n, uint64_t* d_checksums, uint32_t* d_admit_step, void* stream)
{
    if(!s || (n && (!d_arrival || !d_length || !d_ids || !d_checksums)))
        return mmc::fail(MMC_EINVAL, "null argument");
    s->view.arrival = d_arrival;
    s->view.length = d_length;
    s->view.ids = d_ids;
    s->view.checksums = d_checksums;
    s->view.admitStep = d_admit_step;
    s->view.nRequests = n;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(resetKernel, dim3(64), dim3(256), 0, st, s->view);
    if(n)
        hipLaunchKernelGGL(validateKernel, dim3(mmc::gridFor(n, 256) < 1024u ? mmc::gridFor(n, 256) : 1024u),
                           dim3(256), 0, st, s->view, s->cfg.max_tokens);
    u32 firstBad = 0xffffffffu;
    hipError_t err = hipGetLastError();
    if(err == hipSuccess)
        err = hipMemcpyAsync(&firstBad, &s->view.sc->firstBad, sizeof(u32), hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipStreamSynchronize(st);
    if(err != hipSuccess)
        return mmc::hipFail(err, "serve load");
    if(firstBad != 0xffffffffu)
    {
        s->view.nRequests = 0;
        return mmc::fail(MMC_EINVAL, "request " + std::to_string(firstBad) + ": length must be 1.." +
                                         std::to_string(s->cfg.max_tokens) +
                                         " tokens and arrival steps must not decrease");
    }
    hipLaunchKernelGGL(firstAdmitKernel, dim3(1), dim3(64), 0, st, s->view);
    err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "serve reset launch");
}

int mmc_serve_step(mmc_serve* s, void* stream)
{
    if(!s)
        return mmc::fail(MMC_EINVAL, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    withGeometry(
        s->view.heap,
        [&](auto g) {
            hipLaunchKernelGGL((decodeKernel<decltype(g)>), dim3(mmc::gridFor(s->cfg.max_slots, kDecodeThreads)),
                               dim3(kDecodeThreads), 0, st, s->view);
            hipLaunchKernelGGL((completeKernel<decltype(g)>), dim3(kCompleteBlocks + 1u), dim3(256), 0, st, s->view);
        },
        PinnedGeometries{});
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "serve step launch");
}

int mmc_serve_timestamps(mmc_serve* s, uint64_t* d_stamps, uint32_t capacity, uint64_t* clock_hz)
{
    if(!s)
        return mmc::fail(MMC_EINVAL, "null argument");
    s->stamps = capacity ? d_stamps : nullptr;
    s->stampCapacity = d_stamps ? capacity : 0u;
    s->view.stamps = s->stamps;
    s->view.stampCapacity = s->stampCapacity;
    if(clock_hz)
    {
        int khz = 0;
        const hipError_t err = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->heap->device);
        if(err != hipSuccess)
            return mmc::hipFail(err, "hipDeviceGetAttribute(wall clock rate)");
        *clock_hz = static_cast<uint64_t>(khz) * 1000u;
    }
    return MMC_OK;
}

int mmc_serve_stamp(mmc_serve* s, void* stream)
{
    if(!s || !s->stamps)
        return mmc::fail(MMC_EINVAL, "no timestamp buffer attached");
    hipLaunchKernelGGL(stampKernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), s->view.sc, s->stamps,
                       s->stampCapacity);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "stamp launch");
}

int mmc_serve_stamp_count(mmc_serve* s, uint32_t* count, void* stream)
{
    if(!s || !count)
        return mmc::fail(MMC_EINVAL, "null argument");
    ServeScalars h{};
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemcpyAsync(&h, s->view.sc, sizeof(h), hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipStreamSynchronize(st);
    if(err != hipSuccess)
        return mmc::hipFail(err, "stamp count copy");
    *count = h.nStamps;
    return MMC_OK;
}

int mmc_serve_get_stats(mmc_serve* s, mmc_serve_stats* out, void* stream)
{
    if(!s || !out)
        return mmc::fail(MMC_EINVAL, "null argument");
    ServeScalars h{};
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemcpyAsync(&h, s->view.sc, sizeof(h), hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipStreamSynchronize(st);
    if(err != hipSuccess)
        return mmc::hipFail(err, "serve stats copy");
    out->step = h.step > 0u ? h.step - 1u : 0u; // admission rounds include the one at load
    out->admitted = h.counters[cAdmitted];
    out->completed = h.counters[cCompleted];
    out->block_mallocs = h.counters[cMallocs];
    out->block_frees = h.counters[cFrees];
    out->nulls = h.counters[cNulls];
    out->queue_wait_steps = h.counters[cQueueWait];
    out->active = s->cfg.max_slots - h.freeTop;
    out->reserved_bytes = h.reserved;
    out->max_active = h.counters[cMaxActive];
    return MMC_OK;
}

} // extern "C"
