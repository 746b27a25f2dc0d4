// Exhaustive device check of mallocMC::udiv24, the float-reciprocal division
// behind ScatterHeap::pdiv (include/mallocMC/mallocMC_utils.hpp).  pdiv divides
// values up to the pagesize by chunk sizes, segment sizes (32 chunks + 4) and
// size-class bases, all at most the pagesize; so every n <= 65536 against every
// d in 1..65536 covers every pagesize up to 64 KiB completely (4.3G pairs).
// Then 2^28 pairs with n < 2^24 (the bound under which pdiv uses udiv24)
// drawn from splitmix64: exact whenever n / d < 2^22, and min(q, 32) exact
// always (callers clamp quotients to 32 or less).
// Prints the mismatch counts and "udiv24 OK".
#include <mallocMC/mallocMC_utils.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void exhaustiveKernel(unsigned nmax, unsigned long long* mismatches, unsigned long long* first)
{
    const unsigned d = blockIdx.x + 1u; // 1..gridDim.x
    unsigned long long bad = 0;
    for(unsigned n = threadIdx.x; n <= nmax; n += blockDim.x)
    {
        const unsigned q = mallocMC::udiv24(n, d);
        if(q != n / d)
        {
            ++bad;
            atomicMin(first, (static_cast<unsigned long long>(d) << 32) | n);
        }
    }
    if(bad)
        atomicAdd(mismatches, bad);
}

__device__ unsigned long long mix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void sampledKernel(unsigned long long perThread, unsigned long long* exactBad,
                              unsigned long long* clampBad, unsigned long long* exactChecked)
{
    const unsigned long long tid = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x;
    unsigned long long eb = 0, cb = 0, ec = 0;
    for(unsigned long long i = 0; i < perThread; ++i)
    {
        const unsigned long long r = mix((tid << 20) ^ i);
        const unsigned n = static_cast<unsigned>(r) & 0xFFFFFFu;                // < 2^24
        const unsigned shift = static_cast<unsigned>(r >> 32) % 25u;              // divisors of every magnitude
        const unsigned d = 1u + (static_cast<unsigned>(r >> 40) & ((1u << shift) - 1u));
        const unsigned q = mallocMC::udiv24(n, d);
        const unsigned exact = n / d;
        if(exact < (1u << 22))
        {
            ++ec;
            eb += q != exact ? 1u : 0u;
        }
        const unsigned qc = q < 32u ? q : 32u;
        const unsigned ec32 = exact < 32u ? exact : 32u;
        cb += qc != ec32 ? 1u : 0u;
    }
    atomicAdd(exactBad, eb);
    atomicAdd(clampBad, cb);
    atomicAdd(exactChecked, ec);
}

int main()
{
    unsigned long long* d = nullptr;
    if(hipMalloc(&d, 5 * sizeof(unsigned long long)) != hipSuccess)
        return 2;
    unsigned long long h[5] = {0, ~0ull, 0, 0, 0};
    if(hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess)
        return 2;
    const unsigned nmax = 65536u, dmax = 65536u;
    hipLaunchKernelGGL(exhaustiveKernel, dim3(dmax), dim3(256), 0, 0, nmax, d, d + 1);
    const unsigned blocks = 4096u, threads = 256u;
    const unsigned long long perThread = (1ull << 28) / (static_cast<unsigned long long>(blocks) * threads);
    hipLaunchKernelGGL(sampledKernel, dim3(blocks), dim3(threads), 0, 0, perThread, d + 2, d + 3, d + 4);
    if(hipGetLastError() != hipSuccess || hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        return 2;
    (void) hipFree(d);
    std::printf("exhaustive n <= %u, d in 1..%u: %llu mismatches", nmax, dmax, h[0]);
    if(h[0])
        std::printf(" (first d %llu n %llu)", h[1] >> 32, h[1] & 0xFFFFFFFFull);
    std::printf("\nsampled %llu pairs with n < 2^24: %llu exact-range mismatches over %llu, %llu clamped "
                "mismatches\n",
                static_cast<unsigned long long>(blocks) * threads * perThread, h[2], h[4], h[3]);
    const bool ok = h[0] == 0u && h[2] == 0u && h[3] == 0u && h[4] > 0u;
    std::printf("%s\n", ok ? "udiv24 OK" : "udiv24 FAILED");
    return ok ? 0 : 1;
}
