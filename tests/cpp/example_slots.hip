// Device- and host-side getAvailableSlots through the drop-in C++ API, as the
// reference's examples/mallocMC_example03.cpp:76-117 uses them, with the
// XMallocSIMD distribution policy.  Lane 0 mallocs 32 ints, every lane reads
// the slot count of 1-byte slots on the device; the counts must agree with
// the host query.  Prints "example_slots OK" and exits 0 on success.
#include <mallocMC/mallocMC.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>

struct ScatterConfig
{
    static constexpr auto pagesize = 4096;
    static constexpr auto accessblocksize = 64u * 1024u * 1024u;
    static constexpr auto regionsize = 16;
    static constexpr auto wastefactor = 2;
    static constexpr auto resetfreedpages = false;
};

struct XMallocConfig
{
    static constexpr auto pagesize = ScatterConfig::pagesize;
};

using ScatterAllocator = mallocMC::Allocator<
    mallocMC::AccGfx950,
    mallocMC::CreationPolicies::Scatter<ScatterConfig>,
    mallocMC::DistributionPolicies::XMallocSIMD<XMallocConfig>,
    mallocMC::OOMPolicies::ReturnNull,
    mallocMC::ReservePoolPolicies::HipMalloc,
    mallocMC::AlignmentPolicies::Shrink<>>;

__device__ int* arA;

__global__ void exampleKernel(ScatterAllocator::AllocatorHandle h, unsigned* slotsOut, int* ok)
{
    mallocMC::AccGfx950 acc;
    const unsigned id = threadIdx.x;
    if(id == 0)
        arA = static_cast<int*>(h.malloc(acc, sizeof(int) * 32));
    __syncthreads();
    const unsigned slots = h.getAvailableSlots(acc, 1);
    if(arA != nullptr)
        arA[id] = static_cast<int>(id);
    else
        atomicAnd(ok, 0);
    slotsOut[id] = slots;
    __syncthreads();
    if(id == 0)
        h.free(acc, arA);
}

int main()
{
    hipStream_t queue;
    if(hipStreamCreate(&queue) != hipSuccess)
        return 1;
    ScatterAllocator alloc(0, queue, 64u * 1024u * 1024u);
    unsigned* dSlots;
    int* dOk;
    (void) hipMalloc(&dSlots, 32 * sizeof(unsigned));
    (void) hipMalloc(&dOk, sizeof(int));
    const int one = 1;
    (void) hipMemcpy(dOk, &one, sizeof(int), hipMemcpyHostToDevice);
    const unsigned before = alloc.getAvailableSlots(0, queue, 1);
    exampleKernel<<<1, 32, 0, queue>>>(alloc.getAllocatorHandle(), dSlots, dOk);
    unsigned slots[32];
    int ok = 0;
    (void) hipMemcpyAsync(slots, dSlots, sizeof(slots), hipMemcpyDeviceToHost, queue);
    (void) hipMemcpyAsync(&ok, dOk, sizeof(int), hipMemcpyDeviceToHost, queue);
    (void) hipStreamSynchronize(queue);
    const unsigned after = alloc.getAvailableSlots(0, queue, 1);
    std::printf("Slots from Host: %u (before %u); from device: %u\n", after, before, slots[0]);
    // The 128-byte array formats one page for 128-byte chunks, so that page no
    // longer offers 16-byte slots (254 of them at pagesize 4096), neither while
    // the array is live nor after its free (resetfreedpages is off).
    bool same = true;
    for(int i = 1; i < 32; ++i)
        same = same && slots[i] == slots[0];
    if(!ok || !same || slots[0] != before - 254u || after != slots[0])
    {
        std::printf("example_slots FAILED\n");
        return 2;
    }
    std::printf("example_slots OK\n");
    return 0;
}
