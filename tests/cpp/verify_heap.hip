// verify_heap through the drop-in C++ API: the reference's integration
// harness (tests/verify_heap.cpp:86-732) restated for HIP.  Lanes malloc
// 750 x uint64 slots until the heap returns null, write offset+i into each
// element, verify every slot, damage one element and require the verify to
// fail, then free everything and require the slot count to come back.
// Usage: verify_heap [heapMB] [blocks] [threads]; prints "verify_heap OK".
#include <mallocMC/mallocMC.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

constexpr unsigned ELEMS_PER_SLOT = 750;
using allocElem_t = unsigned long long;

struct ScatterConfig
{
    static constexpr auto pagesize = 4096;
    static constexpr auto accessblocksize = 2u * 1024u * 1024u * 1024u;
    static constexpr auto regionsize = 16;
    static constexpr auto wastefactor = 2;
    static constexpr auto resetfreedpages = true;
};

using ScatterAllocator = mallocMC::Allocator<
    mallocMC::AccGfx950,
    mallocMC::CreationPolicies::Scatter<ScatterConfig>,
    mallocMC::DistributionPolicies::Noop,
    mallocMC::OOMPolicies::ReturnNull,
    mallocMC::ReservePoolPolicies::AlpakaBuf<mallocMC::AccGfx950>,
    mallocMC::AlignmentPolicies::Shrink<>>;

__global__ void allocAll(allocElem_t** data, unsigned long long* counter, unsigned long long* globalSum,
                         unsigned long long cap, ScatterAllocator::AllocatorHandle h)
{
    mallocMC::AccGfx950 acc;
    unsigned long long sum = 0;
    while(true)
    {
        allocElem_t* p = static_cast<allocElem_t*>(h.malloc(acc, sizeof(allocElem_t) * ELEMS_PER_SLOT));
        if(p == nullptr)
            break;
        const unsigned long long pos = atomicAdd(counter, 1ull);
        if(pos >= cap)
        {
            h.free(acc, p);
            break;
        }
        const unsigned long long offset = pos * ELEMS_PER_SLOT;
        for(unsigned i = 0; i < ELEMS_PER_SLOT; ++i)
        {
            p[i] = offset + i;
            sum += p[i];
        }
        data[pos] = p;
    }
    atomicAdd(globalSum, sum);
}

__global__ void checkContent(allocElem_t** data, unsigned long long* counter, unsigned long long* globalSum,
                             unsigned long long nSlots, int* correct)
{
    unsigned long long sum = 0;
    while(true)
    {
        const unsigned long long pos = atomicAdd(counter, 1ull);
        if(pos >= nSlots)
            break;
        const unsigned long long offset = pos * ELEMS_PER_SLOT;
        for(unsigned i = 0; i < ELEMS_PER_SLOT; ++i)
        {
            if(data[pos][i] != offset + i)
                atomicAnd(correct, 0);
            sum += data[pos][i];
        }
    }
    atomicAdd(globalSum, sum);
}

__global__ void damageElement(allocElem_t** data)
{
    data[1][0] = static_cast<allocElem_t>(5 * ELEMS_PER_SLOT - 1);
}

__global__ void deallocAll(allocElem_t** data, unsigned long long* counter, unsigned long long nSlots,
                           ScatterAllocator::AllocatorHandle h)
{
    mallocMC::AccGfx950 acc;
    while(true)
    {
        const unsigned long long pos = atomicAdd(counter, 1ull);
        if(pos >= nSlots)
            break;
        h.free(acc, data[pos]);
    }
}

static bool verify(allocElem_t** dData, unsigned long long nSlots, unsigned blocks, unsigned threads,
                   hipStream_t q)
{
    unsigned long long* dCounter;
    unsigned long long* dSum;
    int* dCorrect;
    (void) hipMalloc(&dCounter, 8);
    (void) hipMalloc(&dSum, 8);
    (void) hipMalloc(&dCorrect, 4);
    (void) hipMemsetAsync(dCounter, 0, 8, q);
    (void) hipMemsetAsync(dSum, 0, 8, q);
    const int one = 1;
    (void) hipMemcpyAsync(dCorrect, &one, 4, hipMemcpyHostToDevice, q);
    checkContent<<<blocks, threads, 0, q>>>(dData, dCounter, dSum, nSlots, dCorrect);
    int correct = 0;
    (void) hipMemcpyAsync(&correct, dCorrect, 4, hipMemcpyDeviceToHost, q);
    (void) hipStreamSynchronize(q);
    (void) hipFree(dCounter);
    (void) hipFree(dSum);
    (void) hipFree(dCorrect);
    return correct != 0;
}

int main(int argc, char** argv)
{
    const size_t heapMB = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const unsigned blocks = argc > 2 ? static_cast<unsigned>(std::atoi(argv[2])) : 64;
    const unsigned threads = argc > 3 ? static_cast<unsigned>(std::atoi(argv[3])) : 128;
    const size_t heapSize = heapMB * 1024u * 1024u;
    const size_t slotSize = sizeof(allocElem_t) * ELEMS_PER_SLOT;
    const size_t nPointers = (heapSize + slotSize - 1) / slotSize;
    hipStream_t q;
    if(hipStreamCreate(&q) != hipSuccess)
        return 1;
    bool correct = true;
    unsigned long long used = 0;
    {
        ScatterAllocator alloc(0, q, heapSize);
        const unsigned slots0 = alloc.getAvailableSlots(0, q, slotSize);
        allocElem_t** dData;
        unsigned long long* dCounter;
        unsigned long long* dSum;
        (void) hipMalloc(&dData, sizeof(allocElem_t*) * nPointers);
        (void) hipMalloc(&dCounter, 8);
        (void) hipMalloc(&dSum, 8);
        (void) hipMemsetAsync(dCounter, 0, 8, q);
        (void) hipMemsetAsync(dSum, 0, 8, q);
        allocAll<<<blocks, threads, 0, q>>>(dData, dCounter, dSum, nPointers, alloc.getAllocatorHandle());
        (void) hipMemcpyAsync(&used, dCounter, 8, hipMemcpyDeviceToHost, q);
        (void) hipStreamSynchronize(q);
        if(used > nPointers)
            used = nPointers;
        std::printf("allocated elements:    %llu (%.2f%%)\n", used, 100.0 * used / (heapSize / slotSize));
        correct = correct && used > 0 && verify(dData, used, blocks, threads, q);
        damageElement<<<1, 1, 0, q>>>(dData);
        correct = correct && !verify(dData, used, blocks, threads, q); // must detect the damage
        (void) hipMemsetAsync(dCounter, 0, 8, q);
        deallocAll<<<blocks, threads, 0, q>>>(dData, dCounter, used, alloc.getAllocatorHandle());
        (void) hipStreamSynchronize(q);
        const unsigned slots1 = alloc.getAvailableSlots(0, q, slotSize);
        std::printf("slots of %zu B: %u before, %u after free-all\n", slotSize, slots0, slots1);
        correct = correct && slots1 == slots0;
        (void) hipFree(dData);
        (void) hipFree(dCounter);
        (void) hipFree(dSum);
    }
    std::printf(correct ? "verify_heap OK\n" : "verify_heap FAILED\n");
    return correct ? 0 : 1;
}
