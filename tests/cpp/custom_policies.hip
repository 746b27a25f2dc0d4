// User-written policies plugged into the drop-in headers through the policy
// concept of the reference's DeviceAllocator (device_allocator.hpp:55-107):
//   - RoundUp64, a distribution policy with only the concept members
//     (constructor from acc, collect(acc, bytes), distribute(acc, mem),
//     classname) that rounds every request up to 64 bytes;
//   - CountingNull, an OOM policy (static handleOOM, classname) that counts
//     the failed requests in device memory and returns null.
// Checks, on an 8 MiB heap of 4 KiB pages:
//   1. 4096 threads malloc(16) through the handle; every block is distinct,
//      at least 64 bytes from its neighbours, and the host and device slot
//      counts for 64-byte chunks drop by exactly 4096 (each request took one
//      64-byte chunk);
//   2. 64 threads malloc(1 MiB); the heap holds fewer than 8 such runs, every
//      null is counted by CountingNull (count == nulls > 0);
//   3. the same OOM policy behind the built-in Noop distribution, where full
//      wavefronts take the wave-cooperative path: 64 lanes malloc(4 KiB) on a
//      heap with fewer pages than lanes; the nulls are counted again;
//   4. frees restore the slot counts; Allocator::info() names both policies;
//   5. a full wavefront asking for 40-page runs: the drop-in serves
//      requests above a page per lane (the reference's page-based path),
//      every run distinct and the pages restored.
// Prints "custom_policies OK".
#include <mallocMC/mallocMC.hpp>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

__device__ unsigned gOomCalls;

struct CountingNull
{
    __device__ static void* handleOOM(void*)
    {
        atomicAdd(&gOomCalls, 1u);
        return nullptr;
    }

    static std::string classname()
    {
        return "CountingNull";
    }
};

struct RoundUp64
{
    template<typename Acc>
    __device__ explicit RoundUp64(const Acc&)
    {
    }

    template<typename Acc>
    __device__ std::uint32_t collect(const Acc&, std::uint32_t bytes) const
    {
        return bytes == 0u ? 0u : (bytes + 63u) & ~63u;
    }

    template<typename Acc>
    __device__ void* distribute(const Acc&, void* mem) const
    {
        return mem;
    }

    static std::string classname()
    {
        return "RoundUp64";
    }
};

struct SmallPages
{
    static constexpr auto pagesize = 4096;
    static constexpr auto accessblocksize = 2u * 1024u * 1024u * 1024u;
    static constexpr auto regionsize = 16;
    static constexpr auto wastefactor = 2;
    static constexpr auto resetfreedpages = true;
};

using Custom = mallocMC::Allocator<
    mallocMC::AccGfx950,
    mallocMC::CreationPolicies::Scatter<SmallPages>,
    RoundUp64,
    CountingNull,
    mallocMC::ReservePoolPolicies::HipMalloc,
    mallocMC::AlignmentPolicies::Shrink<>>;

using BuiltinDistribution = mallocMC::Allocator<
    mallocMC::AccGfx950,
    mallocMC::CreationPolicies::Scatter<SmallPages>,
    mallocMC::DistributionPolicies::Noop,
    CountingNull,
    mallocMC::ReservePoolPolicies::HipMalloc,
    mallocMC::AlignmentPolicies::Shrink<>>;

static_assert(!Custom::DevAllocator::waveCapable, "a policy without a Device type is served per lane");
static_assert(BuiltinDistribution::DevAllocator::waveCapable, "Noop keeps the wave path");

template<class Handle>
__global__ void mallocKernel(Handle h, void** ptrs, unsigned n, unsigned bytes)
{
    mallocMC::AccGfx950 acc;
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    if(tid >= n)
        return;
    void* p = h.malloc(acc, bytes);
    if(p != nullptr)
        *static_cast<unsigned*>(p) = tid;
    ptrs[tid] = p;
}

template<class Handle>
__global__ void freeKernel(Handle h, void** ptrs, unsigned n, unsigned* bad)
{
    mallocMC::AccGfx950 acc;
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    if(tid >= n)
        return;
    void* p = ptrs[tid];
    if(p != nullptr && *static_cast<unsigned*>(p) != tid)
        atomicAdd(bad, 1u);
    h.free(acc, p);
}

template<class Handle>
__global__ void slotsKernel(Handle h, unsigned slotSize, unsigned* out)
{
    mallocMC::AccGfx950 acc;
    const unsigned s = h.getAvailableSlots(acc, slotSize); // all 64 lanes of the one wave
    if(threadIdx.x == 0u)
        *out = s;
}

#define CHECK(cond, ...)                                                                                             \
    do                                                                                                               \
    {                                                                                                                \
        if(!(cond))                                                                                                  \
        {                                                                                                            \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                         \
            std::printf(__VA_ARGS__);                                                                                \
            std::printf("\n");                                                                                       \
            return 1;                                                                                                \
        }                                                                                                            \
    } while(0)

#define HIP(call) CHECK((call) == hipSuccess, "%s", #call)

static unsigned oomCalls()
{
    unsigned v = 0u;
    if(hipMemcpyFromSymbol(&v, HIP_SYMBOL(gOomCalls), sizeof(v)) != hipSuccess)
        return ~0u;
    return v;
}

static int resetOomCalls()
{
    const unsigned z = 0u;
    return hipMemcpyToSymbol(HIP_SYMBOL(gOomCalls), &z, sizeof(z)) == hipSuccess ? 0 : 1;
}

template<class Handle>
static unsigned deviceSlots(Handle h, unsigned slotSize, unsigned* dOut, hipStream_t q)
{
    hipLaunchKernelGGL(slotsKernel<Handle>, dim3(1), dim3(64), 0, q, h, slotSize, dOut);
    unsigned v = ~0u;
    if(hipMemcpyAsync(&v, dOut, sizeof(v), hipMemcpyDeviceToHost, q) != hipSuccess
       || hipStreamSynchronize(q) != hipSuccess)
        return ~0u;
    return v;
}

template<class Alloc>
static int runAndFree(Alloc& a, hipStream_t q, void** dPtrs, unsigned n, unsigned bytes, std::vector<void*>& out,
                      unsigned* dBad)
{
    auto h = a.getAllocatorHandle();
    hipLaunchKernelGGL(mallocKernel<decltype(h)>, dim3((n + 255u) / 256u), dim3(256), 0, q, h, dPtrs, n, bytes);
    HIP(hipGetLastError());
    out.assign(n, nullptr);
    HIP(hipMemcpyAsync(out.data(), dPtrs, n * sizeof(void*), hipMemcpyDeviceToHost, q));
    HIP(hipStreamSynchronize(q));
    return 0;
}

template<class Alloc>
static int freeAll(Alloc& a, hipStream_t q, void** dPtrs, unsigned n, unsigned* dBad)
{
    auto h = a.getAllocatorHandle();
    HIP(hipMemsetAsync(dBad, 0, sizeof(unsigned), q));
    hipLaunchKernelGGL(freeKernel<decltype(h)>, dim3((n + 255u) / 256u), dim3(256), 0, q, h, dPtrs, n, dBad);
    HIP(hipGetLastError());
    unsigned bad = ~0u;
    HIP(hipMemcpyAsync(&bad, dBad, sizeof(bad), hipMemcpyDeviceToHost, q));
    HIP(hipStreamSynchronize(q));
    CHECK(bad == 0u, "%u blocks lost their contents", bad);
    return 0;
}

int main()
{
    const int dev = 0;
    HIP(hipSetDevice(dev));
    hipStream_t q;
    HIP(hipStreamCreate(&q));
    constexpr unsigned kMax = 4096u;
    void** dPtrs = nullptr;
    unsigned* dBad = nullptr;
    unsigned* dOut = nullptr;
    HIP(hipMalloc(&dPtrs, kMax * sizeof(void*)));
    HIP(hipMalloc(&dBad, sizeof(unsigned)));
    HIP(hipMalloc(&dOut, sizeof(unsigned)));
    std::vector<void*> ptrs;

    {
        Custom a(dev, q, 8u * 1024u * 1024u);
        const std::string info = Custom::info("\n");
        CHECK(info.find("RoundUp64") != std::string::npos && info.find("CountingNull") != std::string::npos,
              "info() misses the user policies:\n%s", info.c_str());
        const unsigned s64 = a.getAvailableSlots(dev, q, 64);
        const unsigned s16 = a.getAvailableSlots(dev, q, 16);
        CHECK(s64 > kMax, "heap too small: %u slots of 64 B", s64);
        CHECK(deviceSlots(a.getAllocatorHandle(), 64u, dOut, q) == s64, "device and host slot counts differ");

        // 1. every request rounded up to one 64-byte chunk
        CHECK(resetOomCalls() == 0, "reset");
        if(runAndFree(a, q, dPtrs, kMax, 16u, ptrs, dBad))
            return 1;
        std::vector<char*> sorted;
        for(void* p : ptrs)
        {
            CHECK(p != nullptr, "malloc(16) returned null on an empty heap");
            sorted.push_back(static_cast<char*>(p));
        }
        std::sort(sorted.begin(), sorted.end());
        for(std::size_t i = 1; i < sorted.size(); ++i)
            CHECK(sorted[i] - sorted[i - 1] >= 64, "blocks %zu and %zu are %td bytes apart", i - 1, i,
                  sorted[i] - sorted[i - 1]);
        const unsigned after64 = a.getAvailableSlots(dev, q, 64);
        CHECK(after64 == s64 - kMax, "64-byte slots %u -> %u after %u requests", s64, after64, kMax);
        CHECK(deviceSlots(a.getAllocatorHandle(), 64u, dOut, q) == after64, "device slot count differs");
        CHECK(oomCalls() == 0u, "OOM policy called on an empty heap");
        if(freeAll(a, q, dPtrs, kMax, dBad))
            return 1;
        CHECK(a.getAvailableSlots(dev, q, 64) == s64 && a.getAvailableSlots(dev, q, 16) == s16,
              "slots not restored after the 16-byte requests");

        // 2. runs of 256 pages: fewer than 8 fit, every null is counted
        const unsigned nBig = 64u;
        if(runAndFree(a, q, dPtrs, nBig, 1u << 20, ptrs, dBad))
            return 1;
        unsigned nulls = 0u;
        for(unsigned i = 0; i < nBig; ++i)
            nulls += ptrs[i] == nullptr ? 1u : 0u;
        CHECK(nulls > 0u && nulls < nBig, "%u of %u 1 MiB requests failed", nulls, nBig);
        CHECK(oomCalls() == nulls, "CountingNull saw %u calls for %u nulls", oomCalls(), nulls);
        std::printf("1 MiB requests: %u of %u null, %u OOM calls\n", nulls, nBig, oomCalls());
        if(freeAll(a, q, dPtrs, nBig, dBad))
            return 1;
        CHECK(a.getAvailableSlots(dev, q, 64) == s64, "slots not restored after the 1 MiB requests");
    }

    {
        // 3. the wave-cooperative path calls the same OOM policy: 256 KiB = 64 pages minus metadata
        BuiltinDistribution a(dev, q, 256u * 1024u);
        const unsigned pages = a.getAvailableSlots(dev, q, 4096);
        CHECK(pages > 0u && pages < 64u, "%u free pages", pages);
        CHECK(resetOomCalls() == 0, "reset");
        if(runAndFree(a, q, dPtrs, 64u, 4096u, ptrs, dBad))
            return 1;
        unsigned nulls = 0u;
        for(unsigned i = 0; i < 64u; ++i)
            nulls += ptrs[i] == nullptr ? 1u : 0u;
        CHECK(nulls == 64u - pages, "%u nulls for %u pages", nulls, pages);
        CHECK(oomCalls() == nulls, "CountingNull saw %u calls for %u nulls on the wave path", oomCalls(), nulls);
        std::printf("wave path: %u free pages, %u nulls, %u OOM calls\n", pages, nulls, oomCalls());
        if(freeAll(a, q, dPtrs, 64u, dBad))
            return 1;
        CHECK(a.getAvailableSlots(dev, q, 4096) == pages, "pages not restored");
    }

    {
        // 5. requests above a page through the drop-in: 64 lanes of one
        //    wavefront each ask for 40 pages
        BuiltinDistribution a(dev, q, 16u * 1024u * 1024u);
        const unsigned s4k = a.getAvailableSlots(dev, q, 4096);
        const unsigned bytes = 40u * 4096u - 64u;
        if(runAndFree(a, q, dPtrs, 64u, bytes, ptrs, dBad))
            return 1;
        std::vector<char*> sorted;
        for(void* p : ptrs)
        {
            CHECK(p != nullptr, "a 40-page request failed on a heap of %u free pages", s4k);
            sorted.push_back(static_cast<char*>(p));
        }
        std::sort(sorted.begin(), sorted.end());
        for(std::size_t i = 1; i < sorted.size(); ++i)
            CHECK(sorted[i] - sorted[i - 1] >= static_cast<std::ptrdiff_t>(bytes), "runs %zu and %zu overlap", i - 1,
                  i);
        CHECK(a.getAvailableSlots(dev, q, 4096) == s4k - 64u * 40u, "free pages %u -> %u", s4k,
              a.getAvailableSlots(dev, q, 4096));
        if(freeAll(a, q, dPtrs, 64u, dBad))
            return 1;
        CHECK(a.getAvailableSlots(dev, q, 4096) == s4k, "pages not restored after the runs");
        std::printf("40-page runs: 64 of 64 served\n");
    }

    HIP(hipFree(dPtrs));
    HIP(hipFree(dBad));
    HIP(hipFree(dOut));
    HIP(hipStreamDestroy(q));
    std::printf("custom_policies OK\n");
    return 0;
}
