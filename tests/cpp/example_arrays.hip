// Arrays of arrays through the drop-in C++ API, as the reference's
// examples/mallocMC_example01.cpp:82-210 does: one lane allocates three
// pointer tables, every lane allocates three int arrays, fills two, adds them
// into the third; the sums are checked against the closed form and all memory
// is freed.  Prints "example_arrays OK" and exits 0 on success.
#include <mallocMC/mallocMC.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <numeric>
#include <vector>

struct ScatterHeapConfig
{
    static constexpr auto pagesize = 4096;
    static constexpr auto accessblocksize = 256u * 1024u * 1024u;
    static constexpr auto regionsize = 16;
    static constexpr auto wastefactor = 2;
    static constexpr auto resetfreedpages = true;
};

struct ScatterHashConfig
{
    static constexpr auto hashingK = 38183;
    static constexpr auto hashingDistMP = 17497;
    static constexpr auto hashingDistWP = 1;
    static constexpr auto hashingDistWPRel = 1;
};

struct ShrinkConfig
{
    static constexpr auto dataAlignment = 16;
};

using ScatterAllocator = mallocMC::Allocator<
    mallocMC::AccGfx950,
    mallocMC::CreationPolicies::Scatter<ScatterHeapConfig, ScatterHashConfig>,
    mallocMC::DistributionPolicies::Noop,
    mallocMC::OOMPolicies::ReturnNull,
    mallocMC::ReservePoolPolicies::AlpakaBuf<mallocMC::AccGfx950>,
    mallocMC::AlignmentPolicies::Shrink<ShrinkConfig>>;

__device__ int** arA;
__device__ int** arB;
__device__ int** arC;

__global__ void createArrayPointers(int x, int y, ScatterAllocator::AllocatorHandle h)
{
    mallocMC::AccGfx950 acc;
    arA = static_cast<int**>(h.malloc(acc, sizeof(int*) * x * y));
    arB = static_cast<int**>(h.malloc(acc, sizeof(int*) * x * y));
    arC = static_cast<int**>(h.malloc(acc, sizeof(int*) * x * y));
}

__global__ void fillArrays(int length, ScatterAllocator::AllocatorHandle h, int* failures)
{
    mallocMC::AccGfx950 acc;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    arA[id] = static_cast<int*>(h.malloc(acc, length * sizeof(int)));
    arB[id] = static_cast<int*>(h.malloc(acc, length * sizeof(int)));
    arC[id] = static_cast<int*>(h.malloc(acc, length * sizeof(int)));
    if(!arA[id] || !arB[id] || !arC[id])
    {
        atomicAdd(failures, 1);
        return;
    }
    for(int i = 0; i < length; ++i)
    {
        arA[id][i] = id * length + i;
        arB[id][i] = id * length + i;
    }
}

__global__ void addArrays(int length, long long* sums)
{
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    long long s = 0;
    for(int i = 0; i < length; ++i)
    {
        arC[id][i] = arA[id][i] + arB[id][i];
        s += arC[id][i];
    }
    sums[id] = s;
}

__global__ void freeArrays(ScatterAllocator::AllocatorHandle h)
{
    mallocMC::AccGfx950 acc;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    h.free(acc, arA[id]);
    h.free(acc, arB[id]);
    h.free(acc, arC[id]);
}

__global__ void freeArrayPointers(ScatterAllocator::AllocatorHandle h)
{
    mallocMC::AccGfx950 acc;
    h.free(acc, arA);
    h.free(acc, arB);
    h.free(acc, arC);
}

#define CHECK(x)                                                                                                      \
    do                                                                                                                \
    {                                                                                                                 \
        if((x) != hipSuccess)                                                                                         \
        {                                                                                                             \
            std::printf("HIP error at %s:%d\n", __FILE__, __LINE__);                                                  \
            return 1;                                                                                                 \
        }                                                                                                             \
    } while(0)

int main()
{
    const int length = 100;
    const int block = 64;
    const int grid = 64;
    hipStream_t queue;
    CHECK(hipStreamCreate(&queue));
    ScatterAllocator alloc(0, queue, 256u * 1024u * 1024u);
    std::printf("%s", ScatterAllocator::info("\n").c_str());
    const unsigned slots0 = alloc.getAvailableSlots(0, queue, 1024u * 1024u);
    const unsigned small0 = alloc.getAvailableSlots(0, queue, 400);

    int* dFail;
    long long* dSums;
    CHECK(hipMalloc(&dFail, sizeof(int)));
    CHECK(hipMemset(dFail, 0, sizeof(int)));
    CHECK(hipMalloc(&dSums, sizeof(long long) * block * grid));
    createArrayPointers<<<1, 1, 0, queue>>>(grid, block, alloc.getAllocatorHandle());
    fillArrays<<<grid, block, 0, queue>>>(length, alloc.getAllocatorHandle(), dFail);
    int fails = 0;
    CHECK(hipMemcpyAsync(&fails, dFail, sizeof(int), hipMemcpyDeviceToHost, queue));
    CHECK(hipStreamSynchronize(queue));
    if(fails)
    {
        std::printf("%d lanes got null\n", fails);
        return 2;
    }
    addArrays<<<grid, block, 0, queue>>>(length, dSums);
    std::vector<long long> sums(block * grid);
    CHECK(hipMemcpyAsync(sums.data(), dSums, sizeof(long long) * sums.size(), hipMemcpyDeviceToHost, queue));
    CHECK(hipStreamSynchronize(queue));
    const long long sum = std::accumulate(sums.begin(), sums.end(), 0ll);
    const long long n = static_cast<long long>(block) * grid * length;
    const long long gaussian = n * (n - 1);
    std::printf("The sum of the arrays on GPU is %lld\nThe gaussian sum as comparison: %lld\n", sum, gaussian);
    std::printf("there are %u Slots of size 1MB available\n", alloc.getAvailableSlots(0, queue, 1024u * 1024u));
    freeArrays<<<grid, block, 0, queue>>>(alloc.getAllocatorHandle());
    freeArrayPointers<<<1, 1, 0, queue>>>(alloc.getAllocatorHandle());
    CHECK(hipStreamSynchronize(queue));
    const unsigned slots1 = alloc.getAvailableSlots(0, queue, 1024u * 1024u);
    const unsigned small1 = alloc.getAvailableSlots(0, queue, 400);
    if(sum != gaussian || slots1 != slots0 || small1 != small0)
    {
        std::printf("example_arrays FAILED: sum %lld vs %lld, slots %u/%u vs %u/%u\n", sum, gaussian, slots1, small1,
                    slots0, small0);
        return 3;
    }
    CHECK(hipFree(dFail));
    CHECK(hipFree(dSums));
    std::printf("example_arrays OK\n");
    return 0;
}
