// Per-thread malloc and free through the drop-in C++ API, the way a user
// kernel calls them: 1,048,576 threads each malloc 16 * (1 + splitmix64(seed
// << 32 | tid) mod 256) bytes, write their id at both ends of the block,
// and, in a second kernel, check and free.  Run twice: with every lane of a
// wavefront calling together (the allocator serves the wave cooperatively)
// and with lane 63 of every wavefront skipping the call (the active mask is
// not full, so every lane runs the reference's per-lane algorithm).  Both must
// return valid blocks and restore the slot count, and the wave path must not
// return null; prints both rates and "thread_malloc OK".
// Usage: thread_malloc [threads].
#include <mallocMC/mallocMC.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

struct ScatterConfig
{
    static constexpr auto pagesize = 65536;
    static constexpr auto accessblocksize = 2u * 1024u * 1024u * 1024u;
    static constexpr auto regionsize = 32;
    static constexpr auto wastefactor = 2;
    static constexpr auto resetfreedpages = true;
};

using ScatterAllocator = mallocMC::Allocator<
    mallocMC::AccGfx950,
    mallocMC::CreationPolicies::Scatter<ScatterConfig>,
    mallocMC::DistributionPolicies::Noop,
    mallocMC::OOMPolicies::ReturnNull,
    mallocMC::ReservePoolPolicies::HipMalloc,
    mallocMC::AlignmentPolicies::Shrink<>>;

__device__ unsigned long long mix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ unsigned sizeOf(unsigned tid)
{
    return 16u * (1u + static_cast<unsigned>(mix((1ull << 32) | tid) % 256u));
}

__global__ void allocKernel(ScatterAllocator::AllocatorHandle h, unsigned** ptrs, unsigned n, int skipLane63)
{
    mallocMC::AccGfx950 acc;
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    if(tid >= n || (skipLane63 && (threadIdx.x & 63u) == 63u))
        return;
    const unsigned bytes = sizeOf(tid);
    unsigned* p = static_cast<unsigned*>(h.malloc(acc, bytes));
    if(p != nullptr)
    {
        p[0] = tid;
        p[bytes / 4u - 1u] = tid;
    }
    ptrs[tid] = p;
}

__global__ void freeKernel(ScatterAllocator::AllocatorHandle h, unsigned** ptrs, unsigned n, int skipLane63,
                           unsigned* bad)
{
    mallocMC::AccGfx950 acc;
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    if(tid >= n || (skipLane63 && (threadIdx.x & 63u) == 63u))
        return;
    unsigned* p = ptrs[tid];
    const unsigned bytes = sizeOf(tid);
    if(p == nullptr)
        atomicAdd(&bad[1], 1u);
    else if(p[0] != tid || p[bytes / 4u - 1u] != tid)
        atomicAdd(&bad[0], 1u);
    h.free(acc, p);
}

int main(int argc, char** argv)
{
    const unsigned n = argc > 1 ? static_cast<unsigned>(std::strtoul(argv[1], nullptr, 10)) : (1u << 20);
    hipStream_t q;
    if(hipStreamCreate(&q) != hipSuccess)
        return 1;
    bool ok = true;
    {
        ScatterAllocator alloc(0, q, 4ull * 1024u * 1024u * 1024u);
        const unsigned slots0 = alloc.getAvailableSlots(0, q, 64);
        unsigned** dPtrs;
        unsigned* dBad;
        (void) hipMalloc(&dPtrs, sizeof(unsigned*) * n);
        (void) hipMalloc(&dBad, 2 * sizeof(unsigned));
        hipEvent_t e0, e1, e2;
        (void) hipEventCreate(&e0);
        (void) hipEventCreate(&e1);
        (void) hipEventCreate(&e2);
        for(int skip = 0; skip < 2; ++skip)
        {
            for(int rep = 0; rep < 3; ++rep)
            {
                (void) hipMemsetAsync(dBad, 0, 2 * sizeof(unsigned), q);
                (void) hipEventRecord(e0, q);
                allocKernel<<<(n + 255u) / 256u, 256, 0, q>>>(alloc.getAllocatorHandle(), dPtrs, n, skip);
                (void) hipEventRecord(e1, q);
                freeKernel<<<(n + 255u) / 256u, 256, 0, q>>>(alloc.getAllocatorHandle(), dPtrs, n, skip, dBad);
                (void) hipEventRecord(e2, q);
                unsigned bad[2] = {0, 0};
                (void) hipMemcpyAsync(bad, dBad, sizeof(bad), hipMemcpyDeviceToHost, q);
                if(hipStreamSynchronize(q) != hipSuccess)
                    return 1;
                float msMalloc = 0.f, msFree = 0.f;
                (void) hipEventElapsedTime(&msMalloc, e0, e1);
                (void) hipEventElapsedTime(&msFree, e1, e2);
                const unsigned calls = skip ? n - n / 64u : n;
                const unsigned slots = alloc.getAvailableSlots(0, q, 64);
                std::printf(
                    "%s: %u mallocs %.3f ms, frees %.3f ms -> %.3e mallocs+frees/s, nulls %u, bad %u, slots %s\n",
                    skip ? "partial waves (per-lane path)" : "full waves (wave path)", calls, msMalloc, msFree,
                    2.0 * calls / ((msMalloc + msFree) * 1e-3), bad[1], bad[0],
                    slots == slots0 ? "restored" : "NOT restored");
                ok = ok && bad[0] == 0 && slots == slots0 && (skip || bad[1] == 0);
            }
        }
        (void) hipFree(dPtrs);
        (void) hipFree(dBad);
    }
    std::printf(ok ? "thread_malloc OK\n" : "thread_malloc FAILED\n");
    return ok ? 0 : 1;
}
