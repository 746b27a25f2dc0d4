// C ABI of the MI355X-native ScatterAlloc heap: heap lifecycle, slot
// counting, batch malloc/free and the serialized stream.  Declarations and
// reference citations: include/mallocmc_hip.h.
#include "mallocmc_hip.h"

#include "mallocMC/detail/device_allocator.hpp"
#include "mmc_internal.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <sstream>
#include <string>

using namespace mallocMC;
using namespace mallocMC::detail;

namespace mmc
{
    thread_local std::string g_lastError;

    int fail(int code, const std::string& msg)
    {
        g_lastError = msg;
        return code;
    }

    int hipFail(hipError_t e, const char* what)
    {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        return fail(MMC_EHIP, m);
    }

    ScatterConfig toConfig(const mmc_config& c)
    {
        ScatterConfig s;
        s.pagesize = c.pagesize;
        s.regionsize = c.regionsize;
        s.accessblocksize = c.accessblocksize;
        s.wastefactor = c.wastefactor;
        s.resetfreedpages = c.resetfreedpages ? 1u : 0u;
        s.hashingK = c.hashingK;
        s.hashingDistMP = c.hashingDistMP;
        s.hashingDistWP = c.hashingDistWP;
        s.hashingDistWPRel = c.hashingDistWPRel;
        s.dataAlignment = c.dataAlignment;
        s.flags = (c.flags & MMC_FLAG_FIX_PAGESIZE_LEAK) ? kFixPagesizeLeak : 0u;
        return s;
    }

    void toGeometry(const ScatterGeometry& g, mmc_geometry* o)
    {
        o->numpages = g.numpages;
        o->numregions = g.numregions;
        o->accessblocks = g.accessblocks;
        o->pagesperblock = g.pagesperblock;
        o->hierarchyThreshold = g.hierarchyThreshold;
        o->minChunkSize = g.minChunkSize;
        o->maxOnPageMasks = g.maxOnPageMasks;
        o->clampedBlocks = g.clampedBlocks;
        o->ptesOffset = g.ptesOffset;
        o->regionsOffset = g.regionsOffset;
        o->usedBytes = g.usedBytes;
        o->sideMaskWords = g.sideMaskWords;
        o->reserved0 = 0u;
    }

    unsigned gridFor(u64 n, unsigned threads)
    {
        const u64 b = (n + threads - 1u) / threads;
        return static_cast<unsigned>(b == 0 ? 1 : b);
    }
} // namespace mmc

using mmc::fail;
using mmc::hipFail;

// ------------------------------------------------------------------ kernels
namespace
{
    //! Heap init (Scatter.hpp:1026-1048): PTEs, segment masks and region
    //! fill levels to zero, scalars to their start values.
    __global__ void initKernel(ScatterHeap h)
    {
        const u64 tid = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
        const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
        for(u64 i = tid; i < h.numpages; i += stride)
        {
            h.ptes[3u * i] = 0u;
            h.ptes[3u * i + 1u] = 0u;
            h.ptes[3u * i + 2u] = 0u;
        }
        const u64 maskWords = static_cast<u64>(h.numpages) * h.maskStride;
        for(u64 i = tid; i < maskWords; i += stride)
            h.masks[i] = 0u;
        for(u64 i = tid; i < h.numregions; i += stride)
            h.regions[i] = 0u;
        if(tid == 0)
        {
            h.scalars->firstfreeblock = 0u;
            h.scalars->pagebasedMutex = 0u;
            h.scalars->firstFreePageBased = h.numpages - 1u;
        }
    }

    //! getAvailableSlotsHost's kernel (Scatter.hpp:1245-1257) for slot sizes
    //! below a page: page-strided partial counts, summed per workgroup in LDS
    //! and added with one atomic per workgroup (atomics on one address are
    //! performed one after another: one per wave made this launch several
    //! times longer than its PTE reads).
    constexpr unsigned kSlotsThreads = 1024u;
    constexpr unsigned kSlotsMaxBlocks = 512u;

    __global__ __launch_bounds__(kSlotsThreads) void slotsKernel(ScatterHeap h, u64 slotSize, u32* out)
    {
        __shared__ u32 waveSum[kSlotsThreads / 64u];
        const u32 gid = blockIdx.x * blockDim.x + threadIdx.x;
        const u32 stride = gridDim.x * blockDim.x;
        u32 v = h.availableSlotsPartial(slotSize, gid, stride);
        for(int d = 32; d > 0; d >>= 1)
            v += __shfl_xor(v, d, 64);
        if(hw::laneid() == 0u)
            waveSum[threadIdx.x / 64u] = v;
        __syncthreads();
        if(threadIdx.x == 0u)
        {
            u32 total = 0u;
            for(u32 w = 0; w < kSlotsThreads / 64u; ++w)
                total += waveSum[w];
            if(total != 0u)
                atom::add(out, total);
        }
    }

    //! Slot sizes of a page or more: the reference counts, from the top page
    //! down, floor(len / k) slots per maximal run of free pages
    //! (Scatter.hpp:1196-1217).  Phase 1 summarizes contiguous page blocks.
    __global__ void slotsRunsPhase1(ScatterHeap h, u32 k, u32 pagesPerThread, u32 nThreads, u32* summary)
    {
        const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
        if(t >= nThreads)
            return;
        const u64 a64 = static_cast<u64>(t) * pagesPerThread;
        const u32 a = static_cast<u32>(a64 < h.numpages ? a64 : h.numpages);
        const u64 b64 = a64 + pagesPerThread;
        const u32 b = static_cast<u32>(b64 < h.numpages ? b64 : h.numpages);
        // walk downward from b-1 to a, like the reference
        u32 hiRun = 0, loRun = 0, inner = 0, run = 0;
        bool inHi = true;
        for(u32 p = b; p > a;)
        {
            --p;
            const bool isFree = atom::load(h.pteChunksize(p)) == 0u;
            if(isFree)
                ++run;
            else
            {
                if(inHi)
                {
                    hiRun = run;
                    inHi = false;
                }
                else
                    inner += run / k;
                run = 0;
            }
        }
        if(inHi)
            hiRun = run; // whole block free (or empty)
        else
            loRun = run;
        summary[4u * t + 0u] = hiRun;
        summary[4u * t + 1u] = loRun;
        summary[4u * t + 2u] = inner;
        summary[4u * t + 3u] = (inHi ? 1u : 0u) | ((b - a) << 1);
    }

    //! Phase 2: combine the block summaries from the top block down.
    __global__ void slotsRunsPhase2(u32 k, u32 nThreads, const u32* summary, u32* out)
    {
        if(blockIdx.x != 0 || threadIdx.x != 0)
            return;
        u32 carry = 0, total = 0;
        for(u32 t = nThreads; t > 0;)
        {
            --t;
            const u32 hi = summary[4u * t + 0u];
            const u32 lo = summary[4u * t + 1u];
            const u32 inner = summary[4u * t + 2u];
            const bool full = (summary[4u * t + 3u] & 1u) != 0u;
            if(full)
                carry += hi;
            else
            {
                total += (carry + hi) / k;
                total += inner;
                carry = lo;
            }
        }
        total += carry / k;
        *out = total;
    }

    template<class Dist>
    __global__ void mallocBatchKernel(ScatterHeap h, const u32* sizes, u64* offsets, u64 n)
    {
        const u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
        const u32 bytes = i < n ? sizes[i] : 0u;
        void* p = deviceMalloc<WaveIds, Dist>(h, bytes);
        if(i < n)
            offsets[i] = h.offsetOf(p);
    }

    __global__ void freeBatchKernel(ScatterHeap h, const u64* offsets, u64 n)
    {
        const u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
        if(i >= n)
            return;
        const u64 off = offsets[i];
        if(off != kNullOffset)
            deviceFree<WaveIds>(h, h.pages + off);
    }

    //! One lane replays the request stream with the CPU backend's ids.
    __global__ void serialStreamKernel(ScatterHeap h, const int64_t* ops, u64* out, u64 n)
    {
        if(blockIdx.x != 0 || threadIdx.x != 0)
            return;
        for(u64 i = 0; i < n; ++i)
        {
            const int64_t op = ops[i];
            if(op >= 0)
            {
                void* p = deviceMalloc<SerialIds, NoopDistribution>(h, static_cast<u64>(op));
                out[i] = h.offsetOf(p);
            }
            else
            {
                const u64 idx = static_cast<u64>(-(op + 1));
                if(idx < i && out[idx] != kNullOffset)
                    deviceFree<SerialIds>(h, h.pages + out[idx]);
                out[i] = 0u;
            }
        }
    }
} // namespace

// ---------------------------------------------------------------- host side
namespace mmc
{
    constexpr u32 kRunThreads = 16384u;

    //! `external`: a caller-owned, already aligned pool; else the library
    //! reserves one (AlpakaBuf.hpp:41-46) and aligns it (Shrink.hpp:109-130).
    int initHeap(mmc_heap* heap, u64 bytes, hipStream_t stream, void* external = nullptr)
    {
        const ScatterConfig cfg = toConfig(heap->cfg);
        if(const char* e = validateConfig(cfg))
            return fail(MMC_EINVAL, e);
        hipError_t err = hipGetDevice(&heap->device);
        if(err != hipSuccess)
            return hipFail(err, "hipGetDevice");
        void* raw = external;
        u64 memsize = bytes;
        char* pool = static_cast<char*>(raw);
        if(external)
        {
            if((reinterpret_cast<u64>(external) & (cfg.dataAlignment - 1u)) != 0u)
                return fail(MMC_EINVAL, "pool is not aligned to dataAlignment");
            heap->ownsPool = false;
        }
        else
        {
            err = hipMalloc(&raw, bytes);
            if(err != hipSuccess)
                return hipFail(err, "hipMalloc(pool)");
            heap->ownsPool = true;
            pool = static_cast<char*>(raw);
            const u64 misalign = reinterpret_cast<u64>(pool) & (cfg.dataAlignment - 1u);
            if(misalign != 0u)
            {
                pool += cfg.dataAlignment - misalign;
                memsize -= cfg.dataAlignment + misalign;
            }
        }
        heap->rawPool = heap->ownsPool ? raw : nullptr;
        heap->rawBytes = bytes;
        const ScatterGeometry g = computeGeometry(cfg, memsize);
        if(g.numpages == 0u)
        {
            if(heap->ownsPool)
                (void) hipFree(raw);
            heap->rawPool = nullptr;
            return fail(MMC_EINVAL, "heap too small: the geometry has zero pages");
        }
        if(((reinterpret_cast<u64>(pool) + g.ptesOffset) & 7u) != 0u)
        {
            // chunk size and count of a PTE are updated as one 64-bit word
            if(heap->ownsPool)
                (void) hipFree(raw);
            heap->rawPool = nullptr;
            return fail(MMC_EINVAL, "the page table must start 8-byte aligned: use a pagesize that is a multiple "
                                    "of 8 and a pool aligned to 8 bytes");
        }
        err = hipMalloc(&heap->scalars, sizeof(ScatterScalars));
        if(err == hipSuccess)
            err = hipMalloc(&heap->scratch, sizeof(u32) * (4u * kRunThreads + 64u));
        if(err == hipSuccess)
            err = hipMalloc(&heap->masks, sizeof(u32) * static_cast<u64>(g.numpages) * g.sideMaskWords);
        if(err != hipSuccess)
        {
            if(heap->ownsPool)
                (void) hipFree(raw);
            heap->rawPool = nullptr;
            if(heap->scalars)
                (void) hipFree(heap->scalars);
            if(heap->scratch)
                (void) hipFree(heap->scratch);
            heap->scalars = nullptr;
            heap->scratch = nullptr;
            heap->masks = nullptr;
            return hipFail(err, "hipMalloc(heap metadata)");
        }
        // slot-count output and the persistent-kernel work cursor start at zero
        err = hipMemsetAsync(heap->scratch + 4u * kRunThreads, 0, sizeof(u32) * 64u, stream);
        if(err != hipSuccess)
        {
            if(heap->ownsPool)
                (void) hipFree(raw);
            heap->rawPool = nullptr;
            return hipFail(err, "hipMemsetAsync(heap scratch)");
        }
        heap->pool = pool;
        heap->memsize = memsize;
        heap->geom = g;
        heap->view = makeHeapView(cfg, g, pool, memsize, heap->scalars, heap->masks);
        // init kernel: 12 B PTE + 4 B per side mask word per page, HBM-bound; fill the chip
        const unsigned threads = 256;
        u64 blocks64 = (g.numpages + threads - 1u) / threads;
        const unsigned blocks = static_cast<unsigned>(blocks64 < 8192u ? (blocks64 ? blocks64 : 1u) : 8192u);
        hipLaunchKernelGGL(initKernel, dim3(blocks), dim3(threads), 0, stream, heap->view);
        err = hipGetLastError();
        if(err != hipSuccess)
            return hipFail(err, "initKernel launch");
        return MMC_OK;
    }

    void releaseHeap(mmc_heap* heap)
    {
        if(heap->rawPool)
            (void) hipFree(heap->rawPool);
        if(heap->scalars)
            (void) hipFree(heap->scalars);
        if(heap->scratch)
            (void) hipFree(heap->scratch);
        if(heap->masks)
            (void) hipFree(heap->masks);
        heap->rawPool = nullptr;
        heap->scalars = nullptr;
        heap->scratch = nullptr;
        heap->masks = nullptr;
        heap->pool = nullptr;
        heap->memsize = 0;
    }

    int slotsAsync(mmc_heap* heap, u64 slotSize, u32* dOut, hipStream_t stream)
    {
        // Allocator::getAvailableSlots pads first (allocator.hpp:227-233)
        const u64 padded = heap->view.pad(static_cast<u32>(slotSize));
        hipError_t err = hipMemsetAsync(dOut, 0, sizeof(u32), stream);
        if(err != hipSuccess)
            return hipFail(err, "hipMemsetAsync");
        const ScatterHeap& h = heap->view;
        if(padded < h.pagesize)
        {
            u64 blocks = (h.numpages + kSlotsThreads - 1u) / kSlotsThreads;
            blocks = blocks < kSlotsMaxBlocks ? (blocks ? blocks : 1u) : kSlotsMaxBlocks;
            hipLaunchKernelGGL(slotsKernel, dim3(static_cast<unsigned>(blocks)), dim3(kSlotsThreads), 0, stream, h,
                               padded, dOut);
        }
        else
        {
            const u32 k = divup(static_cast<u32>(padded), h.pagesize);
            const u32 per = divup(h.numpages, kRunThreads);
            hipLaunchKernelGGL(slotsRunsPhase1, dim3(kRunThreads / 256u), dim3(256), 0, stream, h, k, per, kRunThreads,
                               heap->scratch);
            hipLaunchKernelGGL(slotsRunsPhase2, dim3(1), dim3(64), 0, stream, k, kRunThreads, heap->scratch, dOut);
        }
        err = hipGetLastError();
        if(err != hipSuccess)
            return hipFail(err, "slots kernel launch");
        return MMC_OK;
    }

    //! Work cursor of the persistent kernels (two words, self-resetting).
    u32* heapCursor(mmc_heap* heap)
    {
        return heap->scratch + 4u * kRunThreads + 16u;
    }

    std::string infoString(const mmc_heap* heap)
    {
        const mmc_config& c = heap->cfg;
        const ScatterGeometry& g = heap->geom;
        std::ostringstream ss;
        // Allocator::info (allocator.hpp:214-223) with Scatter::classname (Scatter.hpp:1354-1370)
        ss << "CreationPolicy:      Scatter[pagesize=" << c.pagesize << ",accessblocksize=" << c.accessblocksize
           << ",regionsize=" << c.regionsize << ",wastefactor=" << c.wastefactor
           << ",resetfreedpages=" << (c.resetfreedpages ? 1 : 0) << ",minChunkSize=" << g.minChunkSize
           << ",HierarchyThreshold=" << g.hierarchyThreshold << ",hashingK=" << c.hashingK
           << ",hashingDistMP=" << c.hashingDistMP << ",hashingDistWP=" << c.hashingDistWP
           << ",hashingDistWPRel=" << c.hashingDistWPRel << "]\n";
        if(c.distribution == MMC_DIST_XMALLOCSIMD)
            ss << "DistributionPolicy:  XMallocSIMD[" << c.pagesize << "]\n";
        else
            ss << "DistributionPolicy:  Noop\n";
        ss << "OOMPolicy:           ReturnNull\n";
        ss << "ReservePoolPolicy:   HipMalloc\n";
        if(c.dataAlignment <= 1u)
            ss << "AlignmentPolicy:     Noop\n";
        else
            ss << "AlignmentPolicy:     Shrink[" << c.dataAlignment << "]\n";
        return ss.str();
    }
} // namespace mmc

// -------------------------------------------------------------------- C ABI
extern "C" {

const char* mmc_last_error(void)
{
    return mmc::g_lastError.c_str();
}

const char* mmc_version(void)
{
    return "mallocmc_hip 0.1.0 (gfx950; Scatter API of mallocMC 2.6.0)";
}

int mmc_default_config(mmc_config* cfg)
{
    if(!cfg)
        return fail(MMC_EINVAL, "null config");
    std::memset(cfg, 0, sizeof(*cfg));
    const ScatterConfig d;
    cfg->pagesize = d.pagesize;
    cfg->regionsize = d.regionsize;
    cfg->accessblocksize = d.accessblocksize;
    cfg->wastefactor = d.wastefactor;
    cfg->resetfreedpages = d.resetfreedpages;
    cfg->hashingK = d.hashingK;
    cfg->hashingDistMP = d.hashingDistMP;
    cfg->hashingDistWP = d.hashingDistWP;
    cfg->hashingDistWPRel = d.hashingDistWPRel;
    cfg->dataAlignment = d.dataAlignment;
    cfg->distribution = MMC_DIST_NOOP;
    cfg->flags = 0u;
    return MMC_OK;
}

int mmc_compute_geometry(const mmc_config* cfg, uint64_t bytes, mmc_geometry* out)
{
    if(!cfg || !out)
        return fail(MMC_EINVAL, "null argument");
    const ScatterConfig c = mmc::toConfig(*cfg);
    if(const char* e = validateConfig(c))
        return fail(MMC_EINVAL, e);
    mmc::toGeometry(computeGeometry(c, bytes), out);
    return MMC_OK;
}

int mmc_create(const mmc_config* cfg, uint64_t bytes, void* stream, mmc_heap** out)
{
    if(!cfg || !out)
        return fail(MMC_EINVAL, "null argument");
    *out = nullptr;
    if(cfg->distribution > MMC_DIST_XMALLOCSIMD)
        return fail(MMC_EINVAL, "unknown distribution policy");
    mmc_heap* heap = new(std::nothrow) mmc_heap();
    if(!heap)
        return fail(MMC_ENOMEM, "host allocation failed");
    heap->cfg = *cfg;
    const int rc = mmc::initHeap(heap, bytes, static_cast<hipStream_t>(stream));
    if(rc != MMC_OK)
    {
        mmc::releaseHeap(heap);
        delete heap;
        return rc;
    }
    *out = heap;
    return MMC_OK;
}

int mmc_create_in_pool(const mmc_config* cfg, void* pool, uint64_t bytes, void* stream, mmc_heap** out)
{
    if(!cfg || !out || (!pool && bytes != 0))
        return fail(MMC_EINVAL, "Scatter policy cannot use nullptr for non-empty memory pools");
    *out = nullptr;
    if(cfg->distribution > MMC_DIST_XMALLOCSIMD)
        return fail(MMC_EINVAL, "unknown distribution policy");
    mmc_heap* heap = new(std::nothrow) mmc_heap();
    if(!heap)
        return fail(MMC_ENOMEM, "host allocation failed");
    heap->cfg = *cfg;
    const int rc = mmc::initHeap(heap, bytes, static_cast<hipStream_t>(stream), pool);
    if(rc != MMC_OK)
    {
        mmc::releaseHeap(heap);
        delete heap;
        return rc;
    }
    *out = heap;
    return MMC_OK;
}

int mmc_destroy(mmc_heap* heap)
{
    if(!heap)
        return MMC_OK;
    (void) hipDeviceSynchronize();
    mmc::releaseHeap(heap);
    delete heap;
    return MMC_OK;
}

int mmc_destructive_resize(mmc_heap* heap, uint64_t bytes, void* stream)
{
    if(!heap)
        return fail(MMC_EINVAL, "null heap");
    if(!heap->ownsPool)
        return fail(MMC_EINVAL, "the pool of this heap is caller-owned: destroy it and create a new one");
    (void) hipDeviceSynchronize();
    mmc::releaseHeap(heap);
    return mmc::initHeap(heap, bytes, static_cast<hipStream_t>(stream));
}

int mmc_heap_location(const mmc_heap* heap, void** pool, uint64_t* size)
{
    if(!heap || !pool || !size)
        return fail(MMC_EINVAL, "null argument");
    *pool = heap->pool;
    *size = heap->memsize;
    return MMC_OK;
}

int mmc_get_config(const mmc_heap* heap, mmc_config* out)
{
    if(!heap || !out)
        return fail(MMC_EINVAL, "null argument");
    *out = heap->cfg;
    return MMC_OK;
}

int mmc_get_geometry(const mmc_heap* heap, mmc_geometry* out)
{
    if(!heap || !out)
        return fail(MMC_EINVAL, "null argument");
    mmc::toGeometry(heap->geom, out);
    return MMC_OK;
}

int mmc_info(const mmc_heap* heap, char* buf, uint64_t len)
{
    if(!heap || !buf || len == 0)
        return fail(MMC_EINVAL, "null argument");
    const std::string s = mmc::infoString(heap);
    const size_t n = s.size() < len - 1 ? s.size() : static_cast<size_t>(len - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
    return MMC_OK;
}

int mmc_available_slots_async(mmc_heap* heap, uint64_t slot_size, uint32_t* d_out, void* stream)
{
    if(!heap || !d_out)
        return fail(MMC_EINVAL, "null argument");
    return mmc::slotsAsync(heap, slot_size, d_out, static_cast<hipStream_t>(stream));
}

int mmc_available_slots(mmc_heap* heap, uint64_t slot_size, void* stream, uint32_t* out)
{
    if(!heap || !out)
        return fail(MMC_EINVAL, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    u32* dOut = heap->scratch + 4u * mmc::kRunThreads;
    int rc = mmc::slotsAsync(heap, slot_size, dOut, s);
    if(rc != MMC_OK)
        return rc;
    hipError_t err = hipMemcpyAsync(out, dOut, sizeof(u32), hipMemcpyDeviceToHost, s);
    if(err == hipSuccess)
        err = hipStreamSynchronize(s);
    if(err != hipSuccess)
        return hipFail(err, "getAvailableSlots copy");
    return MMC_OK;
}

int mmc_get_handle(const mmc_heap* heap, void* dst, uint64_t* size)
{
    if(!heap || !size)
        return fail(MMC_EINVAL, "null argument");
    if(!dst || *size < sizeof(ScatterHeap))
    {
        *size = sizeof(ScatterHeap);
        return fail(MMC_EINVAL, "handle buffer too small");
    }
    std::memcpy(dst, &heap->view, sizeof(ScatterHeap));
    *size = sizeof(ScatterHeap);
    return MMC_OK;
}

int mmc_malloc_batch(mmc_heap* heap, const uint32_t* d_sizes, uint64_t* d_offsets, uint64_t n, void* stream)
{
    if(!heap || (n && (!d_sizes || !d_offsets)))
        return fail(MMC_EINVAL, "null argument");
    if(n == 0)
        return MMC_OK;
    if(n > (1ull << 36))
        return fail(MMC_EINVAL, "batch too large");
    const unsigned threads = 256;
    const unsigned blocks = mmc::gridFor(n, threads);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if(heap->cfg.distribution == MMC_DIST_XMALLOCSIMD)
        hipLaunchKernelGGL(mallocBatchKernel<XMallocSIMD64>, dim3(blocks), dim3(threads), 0, s, heap->view, d_sizes,
                           d_offsets, n);
    else
        hipLaunchKernelGGL(mallocBatchKernel<NoopDistribution>, dim3(blocks), dim3(threads), 0, s, heap->view,
                           d_sizes, d_offsets, n);
    const hipError_t err = hipGetLastError();
    if(err != hipSuccess)
        return hipFail(err, "malloc batch launch");
    return MMC_OK;
}

int mmc_free_batch(mmc_heap* heap, const uint64_t* d_offsets, uint64_t n, void* stream)
{
    if(!heap || (n && !d_offsets))
        return fail(MMC_EINVAL, "null argument");
    if(n == 0)
        return MMC_OK;
    const unsigned threads = 256;
    hipLaunchKernelGGL(freeBatchKernel, dim3(mmc::gridFor(n, threads)), dim3(threads), 0,
                       static_cast<hipStream_t>(stream), heap->view, d_offsets, n);
    const hipError_t err = hipGetLastError();
    if(err != hipSuccess)
        return hipFail(err, "free batch launch");
    return MMC_OK;
}

int mmc_serial_stream(mmc_heap* heap, const int64_t* d_ops, uint64_t* d_out, uint64_t n, void* stream)
{
    if(!heap || (n && (!d_ops || !d_out)))
        return fail(MMC_EINVAL, "null argument");
    if(n == 0)
        return MMC_OK;
    hipLaunchKernelGGL(serialStreamKernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), heap->view, d_ops,
                       d_out, n);
    const hipError_t err = hipGetLastError();
    if(err != hipSuccess)
        return hipFail(err, "serial stream launch");
    return MMC_OK;
}

} // extern "C"
