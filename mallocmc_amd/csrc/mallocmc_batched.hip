// Batched entry points: wave-cooperative malloc/free over one heap (plain or
// persistent launch) and the multi-heap step kernels (persistent, or one
// workgroup per 4, 8 or 16 tiles).
// Declarations: include/mallocmc_batched.h.
#include "mallocmc_batched.h"

#include "mallocMC/detail/device_allocator.hpp"
#include "mallocMC/detail/wave_alloc.hpp"
#include "mmc_internal.hpp"

#include <mutex>
#include <type_traits>
#include <string>
#include <vector>

using namespace mallocMC;
using namespace mallocMC::detail;

struct mmc_multi
{
    ScatterHeap* dHeaps = nullptr;
    std::uint32_t nheaps = 0;
    std::uint32_t distribution = MMC_DIST_NOOP;
    //! position in PinnedGeometries of the configuration every heap has
    //! (their count if they have none of them in common)
    std::uint32_t geometry = 0;
    int device = 0;
};

namespace
{
    //! Block partition of the one-request-per-lane kernels (on by default).
    u32 partitionOf(const mmc_tuning* t)
    {
        return t && t->no_partition ? 0u : 1u;
    }

    //! Threads per block of the one-request-per-lane kernels: 256, 512 or
    //! 1024 (0 = 1024).
    u32 blockOf(const mmc_tuning* t)
    {
        const u32 b = t ? t->block_threads : 0u;
        return b == 0u || b >= 1024u ? 1024u : (b >= 512u ? 512u : 256u);
    }

    WaveTuning toTuning(const mmc_tuning* t)
    {
        WaveTuning w;
        if(t)
        {
            w.aggregation = t->aggregation == 0u ? 64u : (t->aggregation > 64u ? 64u : t->aggregation);
            w.xcdSlices = t->xcd_slices;
            w.classSplit = t->class_split > 3u ? 3u : t->class_split;
            w.flatFast = t->flat_single ? 0u : 1u;
            w.fastForward = t->no_skip ? 0u : 1u;
            w.freshFast = t->fresh_single ? 0u : 1u;
        }
        return w;
    }

    //! Size bucket of a request for the block partition: eight buckets per
    //! octave of (bytes - 1), so that the size classes of a wave malloc
    //! (octaves of the request cut into up to eight parts) are unions of
    //! buckets when the octave base is a power of two.  Larger sizes get
    //! lower buckets: a wave serves its largest class first, and a later,
    //! smaller request can then take a chunk of a page formatted for a larger
    //! size (up to wastefactor times its own) instead of formatting a page of
    //! its own.  Requests above a page and empty lanes get the last buckets.
    constexpr u32 kSizeBuckets = 256u;

    __device__ u32 sizeBucket(u32 bytes, u32 pagesize)
    {
        if(bytes == 0u)
            return kSizeBuckets - 1u;
        if(bytes > pagesize)
            return kSizeBuckets - 2u;
        const u32 x = bytes - 1u;
        if(x < 8u)
            return 223u - x;
        const u32 e = 31u - static_cast<u32>(__builtin_clz(x)); // 3..29
        return 223u - (8u * (e - 2u) + ((x >> (e - 3u)) & 7u));  // 0..215
    }

    //! Sorts the B x M values of a block (B = 256, 512 or 1024 threads, M
    //! values per thread) by an 8-bit key into LDS (counting sort,
    //! unstable): thread t holds values t + j B (j < M) on entry; on return
    //! sortedValue[p] is the value at sorted position p and sortedSlot[p] the
    //! block slot it came from.  Three barriers: after clearing the bucket
    //! counts, after counting, after scattering.  Between the last two,
    //! every wave scans the 256 counts itself and writes the same bucket
    //! starts to LDS (a wave reads only after its own writes, and every wave
    //! writes the same values), instead of four waves scanning and two more
    //! barriers to share their totals.
    template<class T, u32 B, u32 M>
    __device__ void sortByKeyToLds(const u32 (&key)[M], const T (&value)[M], T* sortedValue, u32* sortedSlot)
    {
        static_assert(B % 256u == 0u && B <= 1024u, "a block of 256, 512 or 1024 threads");
        static_assert(kSizeBuckets == 4u * 64u, "four buckets per lane in the scan");
        __shared__ u32 count[kSizeBuckets];
        __shared__ u32 start[kSizeBuckets];
        const u32 tid = threadIdx.x;
        const u32 me = hw::laneid();
        if(tid < kSizeBuckets)
            count[tid] = 0u;
        __syncthreads();
        u32 rank[M];
#pragma unroll
        for(u32 j = 0; j < M; ++j)
            rank[j] = atomicAdd(&count[key[j]], 1u);
        __syncthreads();
        // lane l: buckets 4l .. 4l+3, then a scan of the lane sums
        const u32 c0 = count[4u * me], c1 = count[4u * me + 1u], c2 = count[4u * me + 2u], c3 = count[4u * me + 3u];
        const u32 excl = wave::exclusiveSum(c0 + c1 + c2 + c3);
        start[4u * me] = excl;
        start[4u * me + 1u] = excl + c0;
        start[4u * me + 2u] = excl + c0 + c1;
        start[4u * me + 3u] = excl + c0 + c1 + c2;
        // this wave's writes are in LDS before its reads (LDS operations of a
        // wave complete in order; the fence keeps the compiler from moving
        // the reads up)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for(u32 j = 0; j < M; ++j)
        {
            const u32 pos = start[key[j]] + rank[j];
            sortedValue[pos] = value[j];
            sortedSlot[pos] = tid + j * B;
        }
        __syncthreads();
    }

    //! Reorders the B x M values of a block by an 8-bit key
    //! (sortByKeyToLds): lane l of wave w receives, as its value j, the value
    //! at sorted position 64 (M w + j) + l, with the block slot it came from.
    template<class T, u32 B, u32 M>
    __device__ void partitionByKeyMany(const u32 (&key)[M], const T (&value)[M], T (&outValue)[M],
                                       u32 (&outSlot)[M])
    {
        __shared__ T sortedValue[B * M];
        __shared__ u32 sortedSlot[B * M];
        sortByKeyToLds<T, B, M>(key, value, sortedValue, sortedSlot);
        const u32 first = (threadIdx.x - hw::laneid()) * M + hw::laneid(); // 64 M w + l
#pragma unroll
        for(u32 j = 0; j < M; ++j)
        {
            outValue[j] = sortedValue[first + 64u * j];
            outSlot[j] = sortedSlot[first + 64u * j];
        }
    }

    //! One value per thread: thread t receives the value at sorted position t.
    template<class T, u32 B>
    __device__ void partitionByKey(u32 key, T value, T& outValue, u32& outSlot)
    {
        const u32 k[1] = {key};
        const T v[1] = {value};
        T ov[1];
        u32 os[1];
        partitionByKeyMany<T, B, 1>(k, v, ov, os);
        outValue = ov[0];
        outSlot = os[0];
    }

    //! The B requests of a block are served by its B/64 waves; the whole
    //! wave cooperates.  With `partition` the block's requests are first
    //! reordered by size bucket (partitionByKey), and each lane serves the
    //! request its slot received: uniformly mixed sizes put about six size
    //! classes in a wave of 64 arrivals, and every class costs the wave one
    //! scan and one claim; the larger the block, the fewer waves straddle
    //! two classes.  Without it lane i serves request i.
    template<class Dist, u32 B>
    __global__ __launch_bounds__(B) void waveMallocKernel(ScatterHeap h, WaveTuning t, const u32* sizes,
                                                          u64* offsets, u64 n, u32 partition)
    {
        const u64 base = static_cast<u64>(blockIdx.x) * B;
        const u64 i = base + threadIdx.x;
        u32 bytes = i < n ? sizes[i] : 0u;
        u32 slot = threadIdx.x;
        if(partition != 0u)
            partitionByKey<u32, B>(sizeBucket(bytes, h.pagesize), bytes, bytes, slot);
        const WaveAlloc wa{h, t};
        void* p = wa.malloc<Dist>(bytes);
        if(base + slot < n)
            offsets[base + slot] = h.offsetOf(p);
    }

    //! M requests per lane (M = 2): the block's B x M requests are reordered
    //! by size bucket, and each wave serves 64 M of them through
    //! WaveAlloc::mallocMany, so that a size-class group spans up to 64 M
    //! requests and its scans and claims are paid once for all of them.
    //! Geo and Tun: the heap configuration and the tuning the kernel is
    //! built for (PinnedGeometry and DefaultTuning, or RuntimeGeometry and
    //! RuntimeTuning for any).
    template<u32 B, u32 M, class Geo, class Tun>
    __global__ __launch_bounds__(B) void waveMallocManyKernel(ScatterHeap view, WaveTuning tuning, const u32* sizes,
                                                              u64* offsets, u64 n)
    {
        const ScatterHeap h = Geo::apply(view);
        const WaveTuning t = Tun::apply(tuning);
        const u64 base = static_cast<u64>(blockIdx.x) * (B * M);
        u32 key[M], bytes[M], slot[M];
#pragma unroll
        for(u32 j = 0; j < M; ++j)
        {
            const u64 i = base + threadIdx.x + j * B;
            bytes[j] = i < n ? sizes[i] : 0u;
            key[j] = sizeBucket(bytes[j], h.pagesize);
        }
        partitionByKeyMany<u32, B, M>(key, bytes, bytes, slot);
        const WaveAlloc wa{h, t};
        void* p[M];
        wa.mallocMany<M>(bytes, p);
#pragma unroll
        for(u32 j = 0; j < M; ++j)
            if(base + slot[j] < n)
                offsets[base + slot[j]] = h.offsetOf(p[j]);
    }

    //! Two requests per lane, 1,024-thread blocks: each block sorts 4,096
    //! requests by size bucket in LDS and its 16 waves serve them as 32
    //! windows of 128 sorted requests (WaveAlloc::mallocStream), each wave
    //! taking the next window from a counter in LDS until none is left.  A
    //! wave that finishes early takes more windows, so the few windows of the
    //! smallest size classes (several classes each, served one after the
    //! other) no longer hold a block, and the next block's place on the CU,
    //! after its other waves have ended; sorting 4,096 requests instead of
    //! 2,048 also halves the windows that straddle two classes.  Windows go
    //! out smallest sizes first: the multi-class windows start at once, and
    //! the largest sizes, whose claims take the most pages, come last.
    //! Geo and Tun: the heap configuration and the tuning the kernel is built
    //! for (PinnedGeometry and DefaultTuning, or RuntimeGeometry and
    //! RuntimeTuning for any).
    template<class Geo, class Tun>
    __global__ __launch_bounds__(1024) void waveMallocWindowsKernel(ScatterHeap view, WaveTuning tuning,
                                                                   const u32* sizes, u64* offsets, u64 n)
    {
        constexpr u32 B = 1024u;   // threads
        constexpr u32 kSort = 4u;  // requests a thread brings to the sort
        constexpr u32 M = 2u;      // requests per lane in a window
        constexpr u32 kWindow = 64u * M;
        constexpr u32 kWindows = B * kSort / kWindow;
        const ScatterHeap h = Geo::apply(view);
        const WaveTuning t = Tun::apply(tuning);
        __shared__ u32 sortedBytes[B * kSort];
        __shared__ u32 sortedSlot[B * kSort];
        __shared__ u32 windowsTaken;
        const u64 base = static_cast<u64>(blockIdx.x) * (B * kSort);
        if(threadIdx.x == 0u)
            windowsTaken = 0u; // ordered before its first use by the sort's barriers
        u32 key[kSort], bytes[kSort];
#pragma unroll
        for(u32 j = 0; j < kSort; ++j)
        {
            const u64 i = base + threadIdx.x + j * B;
            bytes[j] = i < n ? sizes[i] : 0u;
            key[j] = sizeBucket(bytes[j], h.pagesize);
        }
        sortByKeyToLds<u32, B, kSort>(key, bytes, sortedBytes, sortedSlot);
        struct Windows
        {
            const ScatterHeap& h;
            const u32* sortedBytes;
            const u32* sortedSlot;
            u32* taken;
            u64* offsets;
            u64 base;
            u64 n;
            u32 first; // the current window's first sorted position
            __device__ bool next(u32 (&b)[M])
            {
                u32 k = 0u;
                if(hw::laneid() == 0u)
                    k = atomicAdd(taken, 1u);
                k = wave::uniform(k);
                if(k >= kWindows)
                    return false;
                first = (kWindows - 1u - k) * kWindow; // smallest sizes first
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    b[j] = sortedBytes[first + 64u * j + hw::laneid()];
                return true;
            }
            __device__ void done(void* const (&res)[M])
            {
                // slots are read from LDS here, not kept in registers
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    const u32 slot = sortedSlot[first + 64u * j + hw::laneid()];
                    if(base + slot < n)
                        offsets[base + slot] = h.offsetOf(res[j]);
                }
            }
        };
        Windows src{h, sortedBytes, sortedSlot, &windowsTaken, offsets, base, n, 0u};
        const WaveAlloc wa{h, t};
        wa.mallocStream<M>(src);
    }

    //! With `partition` the block's frees are first reordered by page index
    //! mod 256, so that the frees of one page gather in one wave and share
    //! its atomics.
    template<u32 B>
    __global__ __launch_bounds__(B) void waveFreeKernel(ScatterHeap h, WaveTuning t, const u64* offsets, u64 n,
                                                        u32 partition)
    {
        const u64 i = static_cast<u64>(blockIdx.x) * B + threadIdx.x;
        u64 off = i < n ? offsets[i] : kNullOffset;
        if(partition != 0u)
        {
            const u32 key = off == kNullOffset ? 255u : (h.pageOf(off) & 255u);
            u32 slot = 0u;
            partitionByKey<u64, B>(key, off, off, slot);
        }
        const WaveAlloc wa{h, t};
        wa.free(off == kNullOffset ? nullptr : h.pages + off);
    }

    //! M frees per lane: the block's B x M frees are reordered by page
    //! index mod 256, and each wave frees 64 M of them through
    //! WaveAlloc::freeMany, so that more of a page's frees meet in one group.
    template<u32 B, u32 M, class Geo>
    __global__ __launch_bounds__(B) void waveFreeManyKernel(ScatterHeap view, WaveTuning t, const u64* offsets, u64 n)
    {
        const ScatterHeap h = Geo::apply(view);
        const u64 base = static_cast<u64>(blockIdx.x) * (B * M);
        u32 key[M], slot[M];
        u64 off[M];
#pragma unroll
        for(u32 j = 0; j < M; ++j)
        {
            const u64 i = base + threadIdx.x + j * B;
            off[j] = i < n ? offsets[i] : kNullOffset;
            key[j] = off[j] == kNullOffset ? 255u : (h.pageOf(off[j]) & 255u);
        }
        partitionByKeyMany<u64, B, M>(key, off, off, slot);
        const WaveAlloc wa{h, t};
        void* p[M];
#pragma unroll
        for(u32 j = 0; j < M; ++j)
            p[j] = off[j] == kNullOffset ? nullptr : h.pages + off[j];
        wa.freeMany<M>(p);
    }

    //! Work-queue helpers shared by the persistent kernels: cursor[0] hands out
    //! tiles, cursor[1] counts finished waves; the last wave to finish resets
    //! both so that the next launch (or graph replay) starts from zero.
    MALLOCMC_D u32 nextTile(u32* cursor)
    {
        u32 tile = 0u;
        if(hw::laneid() == 0u)
            tile = atom::add(&cursor[0], 1u);
        return __builtin_amdgcn_readfirstlane(tile);
    }

    MALLOCMC_D bool finishWave(u32* cursor, u32 totalWaves)
    {
        u32 done = 0u;
        if(hw::laneid() == 0u)
            done = atom::add(&cursor[1], 1u);
        done = __builtin_amdgcn_readfirstlane(done);
        if(done == totalWaves - 1u)
        {
            if(hw::laneid() == 0u)
            {
                atom::store(&cursor[0], 0u);
                atom::store(&cursor[1], 0u);
            }
            return true;
        }
        return false;
    }

    template<class Dist>
    __global__ __launch_bounds__(256) void persistentMallocKernel(ScatterHeap h, WaveTuning t, const u32* sizes,
                                                                  u64* offsets, u64 n, u32* cursor)
    {
        const u32 totalWaves = gridDim.x * (blockDim.x / 64u);
        const u64 nTiles = (n + 63u) / 64u;
        const WaveAlloc wa{h, t};
        for(;;)
        {
            const u32 tile = nextTile(cursor);
            if(tile >= nTiles)
                break;
            const u64 i = static_cast<u64>(tile) * 64u + hw::laneid();
            const u32 bytes = i < n ? sizes[i] : 0u;
            void* p = wa.malloc<Dist>(bytes);
            if(i < n)
                offsets[i] = h.offsetOf(p);
        }
        finishWave(cursor, totalWaves);
    }

    __global__ __launch_bounds__(256) void persistentFreeKernel(ScatterHeap h, WaveTuning t, const u64* offsets, u64 n,
                                                                u32* cursor)
    {
        const u32 totalWaves = gridDim.x * (blockDim.x / 64u);
        const u64 nTiles = (n + 63u) / 64u;
        const WaveAlloc wa{h, t};
        for(;;)
        {
            const u32 tile = nextTile(cursor);
            if(tile >= nTiles)
                break;
            const u64 i = static_cast<u64>(tile) * 64u + hw::laneid();
            const u64 off = i < n ? offsets[i] : kNullOffset;
            wa.free(off == kNullOffset ? nullptr : h.pages + off);
        }
        finishWave(cursor, totalWaves);
    }

    //! A wave-uniform copy of a struct read from memory: every 32-bit word
    //! goes through readfirstlane, so the copy lives in scalar registers.
    template<class T>
    MALLOCMC_D T uniformCopy(const T& src)
    {
        static_assert(sizeof(T) % 4u == 0u, "word-sized struct expected");
        T out;
        const u32* s = reinterpret_cast<const u32*>(&src);
        u32* o = reinterpret_cast<u32*>(&out);
#pragma unroll
        for(u32 i = 0; i < sizeof(T) / 4u; ++i)
            o[i] = __builtin_amdgcn_readfirstlane(s[i]);
        return out;
    }

    MALLOCMC_D u64 mix64(u64 z)
    {
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }

    //! One step over many heaps (see mmc_step_desc).
    template<class Dist>
    __global__ __launch_bounds__(256) void multiStepKernel(const ScatterHeap* heaps, mmc_step_desc d, WaveTuning t)
    {
        const u32 totalWaves = gridDim.x * (blockDim.x / 64u);
        const u64 step = *d.d_step;
        const u32 me = hw::laneid();
        u64 nMalloc = 0, nFree = 0, nNull = 0, nTiles = 0;
        for(;;)
        {
            const u32 tile = nextTile(d.d_tile_cursor);
            if(tile >= d.n_tiles)
                break;
            const u32 first = __builtin_amdgcn_readfirstlane(d.d_tiles[tile]);
            if(first == MMC_NO_TILE)
                continue; // padding
            ++nTiles;
            const u32 heapIdx = __builtin_amdgcn_readfirstlane(d.d_slot_heap[first]);
            const u32 s = first + me;
            const bool valid = s < d.n_slots && d.d_slot_heap[s] == heapIdx;
            // the view is the same for the whole wave: keep it in scalar
            // registers instead of 50 vector registers
            const ScatterHeap h = uniformCopy(heaps[heapIdx]);
            const WaveAlloc wa{h, t};
            const u64 r = mix64(d.seed ^ mix64(step * 0x100000001B3ull + s));
            const u32 coin = static_cast<u32>(r % 1000u);
            u64 off = valid ? d.d_offsets[s] : kNullOffset;
            const bool live = off != kNullOffset;
            const bool doFree = valid && live && coin < d.p_free_permille;
            const bool doAlloc = valid && !live && coin < d.p_alloc_permille;
            wa.free(doFree ? h.pages + off : nullptr);
            u32 bytes = 0u;
            if(doAlloc)
            {
                const u32 u = static_cast<u32>(r >> 32);
                u32 k = 0u;
                while(k + 1u < d.n_sizes && u > d.d_size_cdf[k])
                    ++k;
                bytes = d.d_size_table[k];
            }
            void* p = wa.malloc<Dist>(bytes);
            if(doFree)
            {
                d.d_offsets[s] = kNullOffset;
                d.d_sizes[s] = 0u;
                ++nFree;
            }
            if(doAlloc)
            {
                ++nMalloc;
                if(p == nullptr)
                    ++nNull;
                d.d_offsets[s] = h.offsetOf(p);
                d.d_sizes[s] = p == nullptr ? 0u : bytes;
            }
        }
        // wave totals into the counters
        nMalloc = wave::sum(static_cast<u32>(nMalloc));
        nFree = wave::sum(static_cast<u32>(nFree));
        nNull = wave::sum(static_cast<u32>(nNull));
        if(me == 0u)
        {
            if(nMalloc)
                atom::add(&d.d_counters[0], nMalloc);
            if(nFree)
                atom::add(&d.d_counters[1], nFree);
            if(nNull)
                atom::add(&d.d_counters[2], nNull);
            if(nTiles)
                atom::add(&d.d_counters[3], nTiles);
        }
        if(finishWave(d.d_tile_cursor, totalWaves) && me == 0u)
            atom::add(d.d_step, u64{1});
    }

    //! One step over many heaps, one workgroup per B/64 consecutive 64-slot
    //! tiles (no work queue).  The draws of a slot are those of
    //! multiStepKernel.  When all its tiles with slots lie in one heap the
    //! workgroup first reorders its frees by page and its mallocs by size
    //! (partitionByKey), as the one-heap batch kernels do; the barriers all
    //! come before the wave calls, so a slow wave holds up no other.  The
    //! last workgroup to finish resets the finish counter and advances the
    //! step.
    //! Eight waves per SIMD (64 VGPRs, two 1,024-thread workgroups per CU):
    //! the step's waves mostly wait on atomics, and on the 64-heap workload
    //! two workgroups per CU beat one without spills (1.44e9 against 1.09e9
    //! mallocs+frees/s at 1,024 threads; 512 threads: 1.26e9 against 0.94e9
    //! at the compiler's own 82 VGPRs), the spill of about 27 registers
    //! notwithstanding.
    //! Geo: the heap configuration the kernel is built for (PinnedGeometry,
    //! when every heap of the step has it, or RuntimeGeometry for any).
    template<class Dist, u32 B, class Geo>
    __global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(8))) void multiStepBlockKernel(
        const ScatterHeap* heaps, mmc_step_desc d, WaveTuning t, bool partition)
    {
        constexpr u32 W = B / 64u;
        __shared__ u32 sHeap[W];
        __shared__ u32 sFirst[W];
        __shared__ u32 sLast;
        // the workgroup's mallocs, frees, nulls and tiles, added to the
        // global counters once per workgroup: atomics on one address are
        // performed one after another, and one per wave on the same four
        // words (tens of thousands of waves per step) took longer than the
        // step's own work on nearly empty heaps
        __shared__ unsigned long long sCount[4];
        const u64 step = *d.d_step;
        const u32 me = hw::laneid();
        const u32 w = threadIdx.x / 64u;
        const u32 tile = W * blockIdx.x + w;
        // padding entries (MMC_NO_TILE) give a wave no slots of its own; it
        // still serves its share of a partitioned workgroup's requests
        const u32 entry = tile < d.n_tiles ? __builtin_amdgcn_readfirstlane(d.d_tiles[tile]) : MMC_NO_TILE;
        const bool haveTile = entry != MMC_NO_TILE;
        const u32 first = haveTile ? entry : 0u;
        const u32 heapIdx = haveTile ? __builtin_amdgcn_readfirstlane(d.d_slot_heap[first]) : 0xffffffffu;
        if(me == 0u)
        {
            sHeap[w] = heapIdx;
            sFirst[w] = first;
        }
        if(threadIdx.x < 4u)
            sCount[threadIdx.x] = 0u;
        __syncthreads();
        // the workgroup's heap: that of its first tile with slots; the
        // workgroup partitions its requests when all its tiles with slots
        // lie in that heap
        u32 blockHeap = 0xffffffffu;
        bool oneHeap = true;
        for(u32 i = 0; i < W; ++i)
        {
            if(sHeap[i] == 0xffffffffu)
                continue;
            if(blockHeap == 0xffffffffu)
                blockHeap = sHeap[i];
            else
                oneHeap = oneHeap && sHeap[i] == blockHeap;
        }
        const bool blockWide = partition && blockHeap != 0xffffffffu && oneHeap;
        const u32 s = first + me;
        const bool valid = haveTile && s < d.n_slots && d.d_slot_heap[s] == heapIdx;
        // the view is the same for the whole wave: keep it in scalar registers
        // (a workgroup of padding only reads heap 0 and does nothing with it)
        const u32 viewIdx = haveTile ? heapIdx : (blockHeap != 0xffffffffu ? blockHeap : 0u);
        const ScatterHeap h = Geo::apply(uniformCopy(heaps[__builtin_amdgcn_readfirstlane(viewIdx)]));
        const WaveAlloc wa{h, t};
        const u64 r = mix64(d.seed ^ mix64(step * 0x100000001B3ull + s));
        const u32 coin = static_cast<u32>(r % 1000u);
        const u64 off = valid ? d.d_offsets[s] : kNullOffset;
        const bool live = off != kNullOffset;
        const bool doFree = valid && live && coin < d.p_free_permille;
        const bool doAlloc = valid && !live && coin < d.p_alloc_permille;
        u32 bytes = 0u;
        if(doAlloc)
        {
            const u32 u = static_cast<u32>(r >> 32);
            u32 k = 0u;
            while(k + 1u < d.n_sizes && u > d.d_size_cdf[k])
                ++k;
            bytes = d.d_size_table[k];
        }
        u64 nMalloc = 0, nFree = 0, nNull = 0;
        if(doFree)
        {
            d.d_offsets[s] = kNullOffset;
            d.d_sizes[s] = 0u;
            ++nFree;
        }
        u64 freeOff = doFree ? off : kNullOffset;
        u32 slot = threadIdx.x;
        if(blockWide)
        {
            const u32 key = freeOff == kNullOffset ? 255u : (h.pageOf(freeOff) & 255u);
            partitionByKey<u64, B>(key, freeOff, freeOff, slot);
            slot = threadIdx.x;
            partitionByKey<u32, B>(sizeBucket(bytes, h.pagesize), bytes, bytes, slot);
        }
        wa.free(freeOff == kNullOffset ? nullptr : h.pages + freeOff);
        void* p = wa.malloc<Dist>(bytes);
        if(bytes != 0u)
        {
            // the slot whose request this lane served
            const u32 target = blockWide ? sFirst[slot / 64u] + (slot % 64u) : s;
            ++nMalloc;
            if(p == nullptr)
                ++nNull;
            d.d_offsets[target] = h.offsetOf(p);
            d.d_sizes[target] = p == nullptr ? 0u : bytes;
        }
        nMalloc = wave::sum(static_cast<u32>(nMalloc));
        nFree = wave::sum(static_cast<u32>(nFree));
        nNull = wave::sum(static_cast<u32>(nNull));
        if(me == 0u)
        {
            if(nMalloc)
                atomicAdd(&sCount[0], static_cast<unsigned long long>(nMalloc));
            if(nFree)
                atomicAdd(&sCount[1], static_cast<unsigned long long>(nFree));
            if(nNull)
                atomicAdd(&sCount[2], static_cast<unsigned long long>(nNull));
            if(haveTile)
                atomicAdd(&sCount[3], 1ull);
        }
        // the last workgroup to finish advances the step (every workgroup read
        // it at its start) and resets the finish counter for the next launch
        __syncthreads();
        if(threadIdx.x < 4u && sCount[threadIdx.x] != 0u)
            atom::add(&d.d_counters[threadIdx.x], static_cast<u64>(sCount[threadIdx.x]));
        if(threadIdx.x == 0u)
            sLast = atom::add(&d.d_tile_cursor[1], 1u) == gridDim.x - 1u ? 1u : 0u;
        __syncthreads();
        if(sLast && threadIdx.x == 0u)
        {
            atom::store(&d.d_tile_cursor[1], 0u);
            atom::add(d.d_step, u64{1});
        }
    }

    //! CU count of the current device, queried once per device (never inside
    //! a stream capture after the first call).
    int cuCount()
    {
        static int cached[64] = {0};
        int dev = 0;
        if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
            return 256;
        if(cached[dev] == 0)
        {
            int cus = 256;
            if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
                cus = 256;
            cached[dev] = cus;
        }
        return cached[dev];
    }

    //! Request size of lane `lane` in round `it` of a fill: the two-part
    //! uniform mixture of mmc_fill_desc, from splitmix64(seed, lane, it).
    __device__ u32 fillBytes(const mmc_fill_desc& d, u32 lane, u32 it)
    {
        const u64 r = mix64(d.seed ^ mix64((static_cast<u64>(lane) << 24) + it));
        const u32 coin = static_cast<u32>(r % 1000u);
        const u32 u = static_cast<u32>(r >> 32);
        if(coin < d.p_small_permille)
            return d.small_lo + u % (d.small_hi - d.small_lo + 1u);
        return d.large_lo + u % (d.large_hi - d.large_lo + 1u);
    }

    //! Fill until null (see mmc_fill_desc): each lane allocates until the heap
    //! returns null for it; the wave keeps calling the wave malloc until every
    //! lane stopped.
    //! At most five waves per SIMD: near a full heap more concurrent waves
    //! only contend for the same last pages (at six, the 64 GiB fill to the
    //! first null took 1.3 s per refill instead of 1.2 s).
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 5))) void fillKernel(ScatterHeap h,
                                                                                                WaveTuning t,
                                                                                                mmc_fill_desc d)
    {
        const u32 lane = blockIdx.x * blockDim.x + threadIdx.x;
        bool active = lane < d.lanes;
        const WaveAlloc wa{h, t};
        for(u32 it = 0; hw::ballot(active) != 0u; ++it)
        {
            if(d.d_stop)
            {
                u32 stop = 0u;
                if(hw::laneid() == 0u)
                    stop = atom::load(d.d_stop);
                if(__builtin_amdgcn_readfirstlane(stop) != 0u)
                    break;
            }
            u32 bytes = active ? fillBytes(d, lane, it) : 0u;
            void* p = wa.malloc<NoopDistribution>(bytes);
            if(active)
            {
                if(p == nullptr)
                {
                    active = false;
                    if(d.d_stop && bytes != 0u)
                        atom::store(d.d_stop, 1u);
                }
                else
                {
                    const u64 pos = atom::add(d.d_count, u64{1});
                    if(pos < d.cap)
                    {
                        d.d_offsets[pos] = h.offsetOf(p);
                        d.d_sizes[pos] = bytes;
                    }
                    else
                    {
                        h.destroy<WaveIds>(p);
                        active = false;
                    }
                }
            }
        }
    }

    //! The fill with each workgroup's requests reordered by size before
    //! every round (partitionByKey): lane i's request may be served by
    //! another lane, which reports a null back to lane i through LDS.  The
    //! rounds are workgroup-uniform (the partition has barriers): a
    //! workgroup runs while any of its lanes is active and no stop was seen.
    //! Each lane draws the same sizes as in fillKernel.  On the 64 GiB
    //! stress fill at regionsize 16, workgroups of 1024 lanes refilled in
    //! 0.20-0.24 s, of 512 in 0.27-0.30 s and of 256 in 0.44-0.48 s (arrival
    //! order: 1.0-1.2 s); at regionsize 32, 1024 lanes in 0.17-0.20 s.
    //! A workgroup waits at every round for its longest walk (three quarters
    //! of the waves' time on that fill), yet serving several rounds per
    //! barrier, waves pulling sorted tiles, was 5-8x slower: the waves then
    //! all walk at once and contend for the same last free pages.
    template<u32 B>
    __global__ __launch_bounds__(B) void fillPartitionKernel(ScatterHeap h, WaveTuning t, mmc_fill_desc d)
    {
        __shared__ u32 sStopped[B];
        __shared__ u32 sStop;
        const u32 lane = blockIdx.x * blockDim.x + threadIdx.x;
        bool active = lane < d.lanes;
        const WaveAlloc wa{h, t};
        for(u32 it = 0;; ++it)
        {
            if(threadIdx.x == 0u)
                sStop = d.d_stop ? atom::load(d.d_stop) : 0u;
            sStopped[threadIdx.x] = 0u;
            if(__syncthreads_or(active ? 1 : 0) == 0 || sStop != 0u)
                break;
            u32 bytes = active ? fillBytes(d, lane, it) : 0u;
            u32 slot = threadIdx.x;
            partitionByKey<u32, B>(sizeBucket(bytes, h.pagesize), bytes, bytes, slot);
            void* p = wa.malloc<NoopDistribution>(bytes);
            if(bytes != 0u)
            {
                if(p == nullptr)
                {
                    sStopped[slot] = 1u;
                    if(d.d_stop)
                        atom::store(d.d_stop, 1u);
                }
                else
                {
                    const u64 pos = atom::add(d.d_count, u64{1});
                    if(pos < d.cap)
                    {
                        d.d_offsets[pos] = h.offsetOf(p);
                        d.d_sizes[pos] = bytes;
                    }
                    else
                    {
                        h.destroy<WaveIds>(p);
                        sStopped[slot] = 1u;
                    }
                }
            }
            __syncthreads();
            if(sStopped[threadIdx.x] != 0u)
                active = false;
            __syncthreads();
        }
    }

    //! `geometry`: the position in PinnedGeometries of the configuration
    //! every heap of the table has (kernels built for it run with 1,024-thread
    //! workgroups and the Noop distribution), or the list's length.
    template<u32 B>
    void launchMultiBlock(bool xm, const ScatterHeap* heaps, const mmc_step_desc& d, const WaveTuning& t,
                          bool part, u32 geometry, hipStream_t s)
    {
        const unsigned blocks = (d.n_tiles + B / 64u - 1u) / (B / 64u);
        if(xm)
            hipLaunchKernelGGL((multiStepBlockKernel<XMallocSIMD64, B, RuntimeGeometry>), dim3(blocks), dim3(B), 0,
                               s, heaps, d, t, part);
        else if constexpr(B == 1024u)
            withGeometryAt(
                geometry,
                [&](auto g) {
                    hipLaunchKernelGGL((multiStepBlockKernel<NoopDistribution, B, decltype(g)>), dim3(blocks),
                                       dim3(B), 0, s, heaps, d, t, part);
                },
                PinnedGeometries{});
        else
            hipLaunchKernelGGL((multiStepBlockKernel<NoopDistribution, B, RuntimeGeometry>), dim3(blocks), dim3(B),
                               0, s, heaps, d, t, part);
    }

    unsigned persistentBlocks(const mmc_tuning* t)
    {
        const int cus = cuCount();
        const unsigned wpc = (t && t->waves_per_cu) ? t->waves_per_cu : 8u;
        const unsigned blocks = static_cast<unsigned>(cus) * wpc / 4u; // 4 waves per 256-thread block
        return blocks ? blocks : 1u;
    }

    bool descUsable(const mmc_step_desc* d)
    {
        return d && d->d_slot_heap && d->d_tiles && d->d_size_table && d->d_size_cdf && d->d_offsets
               && d->d_sizes && d->d_step && d->d_counters && d->n_sizes != 0u;
    }

    //! One multi-heap step over the device heap table `heaps`, in the form
    //! the tuning selects.
    void launchStep(const ScatterHeap* heaps, bool xm, const mmc_step_desc& d, const mmc_tuning* tuning,
                    u32 geometry, hipStream_t s)
    {
        const WaveTuning t = toTuning(tuning);
        if(!tuning || tuning->persistent)
        {
            const unsigned blocks = persistentBlocks(tuning);
            if(xm)
                hipLaunchKernelGGL(multiStepKernel<XMallocSIMD64>, dim3(blocks), dim3(256), 0, s, heaps, d, t);
            else
                hipLaunchKernelGGL(multiStepKernel<NoopDistribution>, dim3(blocks), dim3(256), 0, s, heaps, d, t);
            return;
        }
        const bool part = partitionOf(tuning);
        switch(blockOf(tuning))
        {
        case 1024u:
            launchMultiBlock<1024>(xm, heaps, d, t, part, geometry, s);
            break;
        case 512u:
            launchMultiBlock<512>(xm, heaps, d, t, part, geometry, s);
            break;
        default:
            launchMultiBlock<256>(xm, heaps, d, t, part, geometry, s);
            break;
        }
    }

    // ------------------------------------------------ mmc_multi_step_static
    //! Device memory of the allocation-free step: the heap table
    //! (MMC_STATIC_MAX_HEAPS views), then the two-word tile cursor on a
    //! 256-byte line of its own.  48.25 KiB per device, held until the
    //! process ends.
    constexpr u64 kStaticTableBytes = MMC_STATIC_MAX_HEAPS * sizeof(ScatterHeap);
    constexpr u64 kStaticScratchBytes = kStaticTableBytes + 256u;
    static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

    //! Heap views one staging launch carries in its kernel argument (3 KiB).
    constexpr u32 kViewsPerLaunch = 16u;
    struct ViewChunk
    {
        ScatterHeap v[kViewsPerLaunch];
    };
    static_assert(sizeof(ViewChunk) <= 3200u, "a staging launch's argument must stay small");
    static_assert(sizeof(ScatterHeap) % 8u == 0u, "views are copied as 64-bit words");

    //! Copies `count` views from the kernel argument into the table; the
    //! launch that stages the first views also zeroes the tile cursor.
    __global__ __launch_bounds__(256) void stageViewsKernel(ViewChunk c, ScatterHeap* dst, u32 count, u32* cursor)
    {
        const u64* src = reinterpret_cast<const u64*>(&c.v[0]);
        u64* out = reinterpret_cast<u64*>(dst);
        const u32 words = count * static_cast<u32>(sizeof(ScatterHeap) / 8u);
        for(u32 i = threadIdx.x; i < words; i += blockDim.x)
            out[i] = src[i];
        if(cursor != nullptr && threadIdx.x < 2u)
            cursor[threadIdx.x] = 0u;
    }

    //! The array's device address per device, looked up once per device with
    //! that device current (the symbol's host-side name is not a device
    //! address).
    std::mutex gScratchLock;
    char* gScratchAddr[64] = {};

    int staticScratch(int dev, char** out)
    {
        if(dev < 0 || dev >= 64)
            return mmc::fail(MMC_EINVAL, "device ordinal out of range");
        std::lock_guard<std::mutex> lock(gScratchLock);
        if(gScratchAddr[dev] == nullptr)
        {
            void* p = nullptr;
            const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if(err != hipSuccess || p == nullptr)
                return mmc::hipFail(err, "hipGetSymbolAddress(scratch)");
            gScratchAddr[dev] = static_cast<char*>(p);
        }
        *out = gScratchAddr[dev];
        return MMC_OK;
    }
} // namespace

namespace mmc
{
    u32* heapCursor(mmc_heap* heap);
}

extern "C" {

int mmc_selftest_pinned_geometry(uint32_t* pinned, uint32_t* wrong)
{
    if(!pinned || !wrong)
        return mmc::fail(MMC_EINVAL, "null output");
    auto view = [](const ScatterConfig& c) {
        const ScatterGeometry g = computeGeometry(c, 1ull << 30);
        return makeHeapView(c, g, nullptr, 1ull << 30, nullptr, nullptr);
    };
    const u32 sizes[2][2] = {{65536u, 32u}, {16384u, 16u}};
    const u32 more[2][2] = {{65536u, 16u}, {16384u, 32u}};
    u32 ok = 0u, bad = 0u;
    for(const auto* list : {sizes, more})
        for(u32 k = 0; k < 2u; ++k)
        {
            ScatterConfig c;
            c.pagesize = list[k][0];
            c.regionsize = list[k][1];
            c.resetfreedpages = 1u;
            // the configuration itself: exactly one pinned geometry
            ok += countMatches(view(c), PinnedGeometries{}) == 1u ? 1u : 0u;
            // one field changed: none
            ScatterConfig w = c;
            w.wastefactor = 3u;
            bad += countMatches(view(w), PinnedGeometries{});
            w = c;
            w.dataAlignment = 32u;
            bad += countMatches(view(w), PinnedGeometries{});
            w = c;
            w.hashingK += 1u;
            bad += countMatches(view(w), PinnedGeometries{});
            w = c;
            w.regionsize = 8u;
            bad += countMatches(view(w), PinnedGeometries{});
            w = c;
            w.pagesize = 4096u;
            bad += countMatches(view(w), PinnedGeometries{});
        }
    // the default tuning: the kernels built for it run for mmc_default_tuning only
    mmc_tuning d;
    mmc_default_tuning(&d);
    ok += DefaultTuning::matches(toTuning(&d)) ? 1u : 0u;
    for(u32 k = 0; k < 6u; ++k)
    {
        mmc_tuning w = d;
        if(k == 0u)
            w.aggregation = 16u;
        else if(k == 1u)
            w.xcd_slices = 8u;
        else if(k == 2u)
            w.class_split = 1u;
        else if(k == 3u)
            w.flat_single = 1u;
        else if(k == 4u)
            w.no_skip = 1u;
        else
            w.fresh_single = 1u;
        bad += DefaultTuning::matches(toTuning(&w)) ? 1u : 0u;
    }
    *pinned = ok;
    *wrong = bad;
    return MMC_OK;
}

int mmc_default_tuning(mmc_tuning* t)
{
    if(!t)
        return mmc::fail(MMC_EINVAL, "null tuning");
    t->aggregation = 64u;
    t->xcd_slices = 0u;
    t->persistent = 0u;
    t->waves_per_cu = 8u;
    t->class_split = 0u;
    t->flat_single = 0u;
    t->no_skip = 0u;
    t->no_partition = 0u;
    t->block_threads = 0u;
    t->fresh_single = 0u;
    t->lane_requests = 0u;
    return MMC_OK;
}

int mmc_wave_malloc_batch(mmc_heap* heap, const uint32_t* d_sizes, uint64_t* d_offsets, uint64_t n,
                          const mmc_tuning* tuning, void* stream)
{
    if(!heap || (n && (!d_sizes || !d_offsets)))
        return mmc::fail(MMC_EINVAL, "null argument");
    if(n == 0)
        return MMC_OK;
    if(n > (1ull << 37))
        return mmc::fail(MMC_EINVAL, "batch too large");
    const WaveTuning t = toTuning(tuning);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool xm = heap->cfg.distribution == MMC_DIST_XMALLOCSIMD;
    if(tuning && tuning->persistent)
    {
        const unsigned blocks = persistentBlocks(tuning);
        u32* cur = mmc::heapCursor(heap);
        if(xm)
            hipLaunchKernelGGL(persistentMallocKernel<XMallocSIMD64>, dim3(blocks), dim3(256), 0, s, heap->view, t,
                               d_sizes, d_offsets, n, cur);
        else
            hipLaunchKernelGGL(persistentMallocKernel<NoopDistribution>, dim3(blocks), dim3(256), 0, s, heap->view,
                               t, d_sizes, d_offsets, n, cur);
    }
    else
    {
        const u32 part = partitionOf(tuning);
        const u32 bs = blockOf(tuning);
        const unsigned blocks = mmc::gridFor(n, bs);
        if(!xm && part != 0u && tuning && tuning->lane_requests >= 4u)
        {
            const unsigned blocks4 = mmc::gridFor(n, 4u * bs);
            if(bs == 1024u)
                hipLaunchKernelGGL((waveMallocManyKernel<1024, 4, RuntimeGeometry, RuntimeTuning>), dim3(blocks4),
                                   dim3(1024), 0, s, heap->view, t, d_sizes, d_offsets, n);
            else if(bs == 512u)
                hipLaunchKernelGGL((waveMallocManyKernel<512, 4, RuntimeGeometry, RuntimeTuning>), dim3(blocks4),
                                   dim3(512), 0, s, heap->view, t, d_sizes, d_offsets, n);
            else
                hipLaunchKernelGGL((waveMallocManyKernel<256, 4, RuntimeGeometry, RuntimeTuning>), dim3(blocks4),
                                   dim3(256), 0, s, heap->view, t, d_sizes, d_offsets, n);
        }
        else if(!xm && part != 0u && tuning && tuning->lane_requests >= 2u)
        {
            // two requests per lane: a block serves 2 x bs of them
            const unsigned blocks2 = mmc::gridFor(n, 2u * bs);
            if(bs == 1024u)
                withGeometry(
                    heap->view,
                    [&](auto g) {
                        using G = decltype(g);
                        const dim3 grid(mmc::gridFor(n, 4096u));
                        // the default tuning is pinned only beside a pinned geometry
                        if constexpr(!std::is_same_v<G, RuntimeGeometry>)
                        {
                            if(DefaultTuning::matches(t))
                            {
                                hipLaunchKernelGGL((waveMallocWindowsKernel<G, DefaultTuning>), grid, dim3(1024), 0,
                                                   s, heap->view, t, d_sizes, d_offsets, n);
                                return;
                            }
                        }
                        hipLaunchKernelGGL((waveMallocWindowsKernel<G, RuntimeTuning>), grid, dim3(1024), 0, s,
                                           heap->view, t, d_sizes, d_offsets, n);
                    },
                    PinnedGeometries{});
            else if(bs == 512u)
                hipLaunchKernelGGL((waveMallocManyKernel<512, 2, RuntimeGeometry, RuntimeTuning>), dim3(blocks2),
                                   dim3(512), 0, s, heap->view, t, d_sizes, d_offsets, n);
            else
                hipLaunchKernelGGL((waveMallocManyKernel<256, 2, RuntimeGeometry, RuntimeTuning>), dim3(blocks2),
                                   dim3(256), 0, s, heap->view, t, d_sizes, d_offsets, n);
        }
        else if(xm)
        {
            if(bs == 1024u)
                hipLaunchKernelGGL((waveMallocKernel<XMallocSIMD64, 1024>), dim3(blocks), dim3(1024), 0, s,
                                   heap->view, t, d_sizes, d_offsets, n, part);
            else if(bs == 512u)
                hipLaunchKernelGGL((waveMallocKernel<XMallocSIMD64, 512>), dim3(blocks), dim3(512), 0, s, heap->view,
                                   t, d_sizes, d_offsets, n, part);
            else
                hipLaunchKernelGGL((waveMallocKernel<XMallocSIMD64, 256>), dim3(blocks), dim3(256), 0, s, heap->view,
                                   t, d_sizes, d_offsets, n, part);
        }
        else
        {
            if(bs == 1024u)
                hipLaunchKernelGGL((waveMallocKernel<NoopDistribution, 1024>), dim3(blocks), dim3(1024), 0, s,
                                   heap->view, t, d_sizes, d_offsets, n, part);
            else if(bs == 512u)
                hipLaunchKernelGGL((waveMallocKernel<NoopDistribution, 512>), dim3(blocks), dim3(512), 0, s,
                                   heap->view, t, d_sizes, d_offsets, n, part);
            else
                hipLaunchKernelGGL((waveMallocKernel<NoopDistribution, 256>), dim3(blocks), dim3(256), 0, s,
                                   heap->view, t, d_sizes, d_offsets, n, part);
        }
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "wave malloc launch");
}

int mmc_wave_free_batch(mmc_heap* heap, const uint64_t* d_offsets, uint64_t n, const mmc_tuning* tuning,
                        void* stream)
{
    if(!heap || (n && !d_offsets))
        return mmc::fail(MMC_EINVAL, "null argument");
    if(n == 0)
        return MMC_OK;
    const WaveTuning t = toTuning(tuning);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if(tuning && tuning->persistent)
        hipLaunchKernelGGL(persistentFreeKernel, dim3(persistentBlocks(tuning)), dim3(256), 0, s, heap->view, t,
                           d_offsets, n, mmc::heapCursor(heap));
    else
    {
        const u32 bs = blockOf(tuning);
        // with lane_requests 2 or 4, two frees per lane (four measured
        // slower: a launch of one workgroup per CU waits for its slowest wave)
        const bool two = tuning && partitionOf(tuning) && tuning->lane_requests >= 2u;
        if(two && bs == 1024u)
            withGeometry(
                heap->view,
                [&](auto g) {
                    hipLaunchKernelGGL((waveFreeManyKernel<1024, 2, decltype(g)>), dim3(mmc::gridFor(n, 2048)),
                                       dim3(1024), 0, s, heap->view, t, d_offsets, n);
                },
                PinnedGeometries{});
        else if(two && bs == 512u)
            hipLaunchKernelGGL((waveFreeManyKernel<512, 2, RuntimeGeometry>), dim3(mmc::gridFor(n, 1024)), dim3(512), 0,
                               s, heap->view, t, d_offsets, n);
        else if(two)
            hipLaunchKernelGGL((waveFreeManyKernel<256, 2, RuntimeGeometry>), dim3(mmc::gridFor(n, 512)), dim3(256), 0,
                               s, heap->view, t, d_offsets, n);
        else if(bs == 1024u)
            hipLaunchKernelGGL(waveFreeKernel<1024>, dim3(mmc::gridFor(n, 1024)), dim3(1024), 0, s, heap->view, t,
                               d_offsets, n, partitionOf(tuning));
        else if(bs == 512u)
            hipLaunchKernelGGL(waveFreeKernel<512>, dim3(mmc::gridFor(n, 512)), dim3(512), 0, s, heap->view, t,
                               d_offsets, n, partitionOf(tuning));
        else
            hipLaunchKernelGGL(waveFreeKernel<256>, dim3(mmc::gridFor(n, 256)), dim3(256), 0, s, heap->view, t,
                               d_offsets, n, partitionOf(tuning));
    }
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "wave free launch");
}

int mmc_wave_fill_until_null(mmc_heap* heap, const mmc_fill_desc* desc, const mmc_tuning* tuning, void* stream)
{
    if(!heap || !desc || !desc->d_offsets || !desc->d_sizes || !desc->d_count || desc->lanes == 0
       || desc->small_hi < desc->small_lo || desc->large_hi < desc->large_lo || desc->p_small_permille > 1000u
       || desc->small_hi - desc->small_lo == 0xffffffffu || desc->large_hi - desc->large_lo == 0xffffffffu)
        return mmc::fail(MMC_EINVAL, "bad fill descriptor");
    const WaveTuning t = toTuning(tuning);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const u32 bs = blockOf(tuning);
    if(partitionOf(tuning) && bs == 1024u)
        hipLaunchKernelGGL((fillPartitionKernel<1024>), dim3(mmc::gridFor(desc->lanes, 1024)), dim3(1024), 0, s,
                           heap->view, t, *desc);
    else if(partitionOf(tuning) && bs == 512u)
        hipLaunchKernelGGL((fillPartitionKernel<512>), dim3(mmc::gridFor(desc->lanes, 512)), dim3(512), 0, s,
                           heap->view, t, *desc);
    else if(partitionOf(tuning))
        hipLaunchKernelGGL((fillPartitionKernel<256>), dim3(mmc::gridFor(desc->lanes, 256)), dim3(256), 0, s,
                           heap->view, t, *desc);
    else
        hipLaunchKernelGGL(fillKernel, dim3(mmc::gridFor(desc->lanes, 256)), dim3(256), 0,
                           static_cast<hipStream_t>(stream), heap->view, t, *desc);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "fill launch");
}

int mmc_multi_create(mmc_heap* const* heaps, uint32_t nheaps, mmc_multi** out)
{
    if(!heaps || nheaps == 0 || !out)
        return mmc::fail(MMC_EINVAL, "bad argument");
    *out = nullptr;
    std::vector<ScatterHeap> views(nheaps);
    const u32 dist = heaps[0]->cfg.distribution;
    for(u32 i = 0; i < nheaps; ++i)
    {
        if(!heaps[i])
            return mmc::fail(MMC_EINVAL, "null heap in group");
        if(heaps[i]->cfg.distribution != dist)
            return mmc::fail(MMC_EINVAL, "heaps of one group must share the distribution policy");
        views[i] = heaps[i]->view;
    }
    mmc_multi* m = new mmc_multi();
    m->nheaps = nheaps;
    m->distribution = dist;
    m->geometry = commonGeometry(views.data(), nheaps, PinnedGeometries{});
    hipError_t err = hipMalloc(&m->dHeaps, sizeof(ScatterHeap) * nheaps);
    if(err == hipSuccess)
        err = hipMemcpy(m->dHeaps, views.data(), sizeof(ScatterHeap) * nheaps, hipMemcpyHostToDevice);
    if(err != hipSuccess)
    {
        if(m->dHeaps)
            (void) hipFree(m->dHeaps);
        delete m;
        return mmc::hipFail(err, "multi-heap table");
    }
    *out = m;
    return MMC_OK;
}

int mmc_multi_destroy(mmc_multi* m)
{
    if(!m)
        return MMC_OK;
    (void) hipDeviceSynchronize();
    if(m->dHeaps)
        (void) hipFree(m->dHeaps);
    delete m;
    return MMC_OK;
}

int mmc_multi_step(mmc_multi* m, const mmc_step_desc* desc, const mmc_tuning* tuning, void* stream)
{
    if(!m || !descUsable(desc) || !desc->d_tile_cursor)
        return mmc::fail(MMC_EINVAL, "bad step descriptor");
    if(desc->n_tiles == 0)
        return MMC_OK;
    launchStep(m->dHeaps, m->distribution == MMC_DIST_XMALLOCSIMD, *desc, tuning, m->geometry,
               static_cast<hipStream_t>(stream));
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "multi step launch");
}

int mmc_multi_step_static(mmc_heap* const* heaps, uint32_t nheaps, const mmc_step_desc* desc,
                          const mmc_tuning* tuning, void* stream)
{
    if(!heaps || !descUsable(desc))
        return mmc::fail(MMC_EINVAL, "bad argument");
    if(nheaps == 0u || nheaps > MMC_STATIC_MAX_HEAPS)
        return mmc::fail(MMC_EINVAL, "mmc_multi_step_static serves 1 to " + std::to_string(MMC_STATIC_MAX_HEAPS)
                                         + " heaps, got " + std::to_string(nheaps));
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if(err != hipSuccess)
        return mmc::hipFail(err, "hipGetDevice");
    const u32 dist = heaps[0] ? heaps[0]->cfg.distribution : 0u;
    for(u32 i = 0; i < nheaps; ++i)
    {
        if(!heaps[i])
            return mmc::fail(MMC_EINVAL, "null heap in group");
        if(heaps[i]->cfg.distribution != dist)
            return mmc::fail(MMC_EINVAL, "heaps of one step must share the distribution policy");
        if(heaps[i]->device != dev)
            return mmc::fail(MMC_EINVAL, "every heap must live on the current device");
    }
    if(desc->n_tiles == 0)
        return MMC_OK;
    char* base = nullptr;
    const int rc = staticScratch(dev, &base);
    if(rc != MMC_OK)
        return rc;
    ScatterHeap* table = reinterpret_cast<ScatterHeap*>(base);
    u32* cursor = reinterpret_cast<u32*>(base + kStaticTableBytes);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the array keeps the last call's table: rewrite this call's views, and
    // start the tile cursor from zero (kernels, not memsets, so a capture
    // records kernel nodes only)
    // the configuration every heap has, among those with kernels of their own
    // (a view matches at most one of them: each pins another pagesize and
    // regionsize)
    u32 geometry = commonGeometry(&heaps[0]->view, 1u, PinnedGeometries{});
    for(u32 i = 1; i < nheaps; ++i)
        if(commonGeometry(&heaps[i]->view, 1u, PinnedGeometries{}) != geometry)
            geometry = geometryCount(PinnedGeometries{});
    for(u32 first = 0; first < nheaps; first += kViewsPerLaunch)
    {
        ViewChunk c;
        const u32 n = nheaps - first < kViewsPerLaunch ? nheaps - first : kViewsPerLaunch;
        for(u32 i = 0; i < n; ++i)
            c.v[i] = heaps[first + i]->view;
        hipLaunchKernelGGL(stageViewsKernel, dim3(1), dim3(256), 0, s, c, table + first, n,
                           first == 0u ? cursor : nullptr);
    }
    mmc_step_desc d = *desc;
    d.d_tile_cursor = cursor;
    launchStep(table, dist == MMC_DIST_XMALLOCSIMD, d, tuning, geometry, s);
    err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "multi step launch");
}

} // extern "C"
