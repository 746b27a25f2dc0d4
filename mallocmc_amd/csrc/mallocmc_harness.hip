// Harness kernels: verify_heap's fill / check / alloc-until-null
// (tests/verify_heap.cpp:241-405) over offset arrays, and the device-atomic
// ceiling and copy microbenchmarks.  Declarations: include/mallocmc_harness.h.
#include "mallocmc_harness.h"

#include "mallocMC/detail/device_allocator.hpp"
#include "mallocMC/detail/wave_alloc.hpp"
#include "mmc_internal.hpp"

using namespace mallocMC;
using namespace mallocMC::detail;

namespace
{
    //! Each wave walks allocations; lanes stride over the words of one
    //! allocation so stores coalesce.
    __global__ void writePatternKernel(ScatterHeap h, const u64* offsets, const u32* sizes, u64 n, u64 salt)
    {
        const u64 wave = (static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x) / 64u;
        const u64 nWaves = (static_cast<u64>(gridDim.x) * blockDim.x) / 64u;
        const u32 lane = hw::laneid();
        for(u64 i = wave; i < n; i += nWaves)
        {
            const u64 off = offsets[i];
            if(off == kNullOffset)
                continue;
            u64* p = reinterpret_cast<u64*>(h.pages + off);
            const u32 words = sizes[i] / 8u;
            const u64 v = salt + i;
            for(u32 w = lane; w < words; w += 64u)
                p[w] = v;
        }
    }

    __global__ __launch_bounds__(256) void verifyPatternKernel(ScatterHeap h, const u64* offsets, const u32* sizes, u64 n, u64 salt,
                                        u64* result)
    {
        const u64 wave = (static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x) / 64u;
        const u64 nWaves = (static_cast<u64>(gridDim.x) * blockDim.x) / 64u;
        const u32 lane = hw::laneid();
        const u64 pageBytes = static_cast<u64>(h.numpages) * h.pagesize;
        u64 bad = 0, misaligned = 0, outside = 0, nulls = 0, checked = 0;
        for(u64 i = wave; i < n; i += nWaves)
        {
            const u64 off = offsets[i];
            const u32 size = sizes[i];
            if(off == kNullOffset)
            {
                if(lane == 0 && size != 0u)
                    ++nulls;
                continue;
            }
            const u32 padded = h.pad(size);
            if(lane == 0)
            {
                ++checked;
                if((off & h.alignMask) != 0u)
                    ++misaligned;
                if(off + padded > pageBytes)
                    ++outside;
            }
            if(off + padded > pageBytes)
                continue;
            const u64* p = reinterpret_cast<const u64*>(h.pages + off);
            const u32 words = size / 8u;
            const u64 v = salt + i;
            for(u32 w = lane; w < words; w += 64u)
                bad += p[w] != v ? 1u : 0u;
        }
        for(int d = 32; d > 0; d >>= 1)
        {
            bad += __shfl_xor(bad, d, 64);
            misaligned += __shfl_xor(misaligned, d, 64);
            outside += __shfl_xor(outside, d, 64);
            nulls += __shfl_xor(nulls, d, 64);
            checked += __shfl_xor(checked, d, 64);
        }
        // workgroup totals first: one atomic per counter and workgroup, not
        // per wave (atomics on one address are performed one after another)
        __shared__ u64 sums[5][4];
        const u32 w = threadIdx.x / 64u;
        if(lane == 0)
        {
            sums[0][w] = bad;
            sums[1][w] = misaligned;
            sums[2][w] = outside;
            sums[3][w] = nulls;
            sums[4][w] = checked;
        }
        __syncthreads();
        if(threadIdx.x < 5u)
        {
            const u64 t = sums[threadIdx.x][0] + sums[threadIdx.x][1] + sums[threadIdx.x][2] + sums[threadIdx.x][3];
            if(t)
                atom::add(&result[threadIdx.x], t);
        }
    }

    __global__ void allocUntilNullKernel(ScatterHeap h, u32 size, u64* offsets, u64 cap, u64* count)
    {
        for(;;)
        {
            void* p = deviceMalloc<WaveIds, NoopDistribution>(h, size);
            if(p == nullptr)
                break;
            const u64 pos = atom::add(count, u64{1});
            if(pos < cap)
                offsets[pos] = h.offsetOf(p);
            else
            {
                deviceFree<WaveIds>(h, p);
                break;
            }
        }
    }

    __global__ void atomicsKernel(u32* buf, u64 words, u32 pattern, u32 iters, u64* sink)
    {
        const u64 gid = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
        const u64 wave = gid / 64u;
        const u32 lane = hw::laneid();
        u64 idx;
        if(pattern == 1u)
            idx = (wave * 32u) % words; // one word per wave, 128 B apart
        else if(pattern == 2u)
            idx = (wave * 32u + (lane & 31u)) % words; // one 128 B line per wave
        else
            idx = (gid * 32u) % words; // one line per lane
        u32 acc = 0u;
        for(u32 it = 0; it < iters; ++it)
        {
            if(pattern == 3u)
            {
                const u32 old = atom::load(&buf[idx]);
                acc += atom::cas(&buf[idx], old, old + 1u);
            }
            else
                acc += atom::add(&buf[idx], 1u);
        }
        if(acc == 0xdeadbeefu)
            atom::add(sink, u64{1});
    }

    __global__ void copyKernel(const uint4* src, uint4* dst, u64 n16)
    {
        const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
        for(u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
            dst[i] = src[i];
    }
} // namespace

namespace
{
    //! Seeded values per lane; the wave's primitives against lane-by-lane loops.
    __global__ __launch_bounds__(64) void waveSelfTestKernel(u64 seed, u32* result)
    {
        using namespace mallocMC::detail;
        const u32 me = hw::laneid();
        u64 z = seed + 0x9E3779B97F4A7C15ull * (static_cast<u64>(blockIdx.x) * 64u + me + 1u);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        // vary the value range by wave so that ties, zeros and all-ones occur
        const u32 shift = blockIdx.x % 33u;
        const u32 v = static_cast<u32>(z) >> shift;
        const u32 small = static_cast<u32>(z >> 40) % 33u; // 0..32, the prefix domain
        u32 bmax = 0u, bmin = 0xffffffffu, bor = 0u, bpre = 0u, vpre = 0u;
        for(u32 l = 0; l < 64u; ++l)
        {
            const u32 o = static_cast<u32>(__builtin_amdgcn_readlane(static_cast<int>(v), static_cast<int>(l)));
            const u32 so = static_cast<u32>(__builtin_amdgcn_readlane(static_cast<int>(small), static_cast<int>(l)));
            bmax = o > bmax ? o : bmax;
            bmin = o < bmin ? o : bmin;
            bor |= o;
            bpre += l < me ? so : 0u;
            vpre += l < me ? o : 0u; // modulo 2^32
        }
        const u32 m0 = hw::ballot(wave::max(v) != bmax) != 0u ? 1u : 0u;
        const u32 m1 = hw::ballot(wave::min(v) != bmin) != 0u ? 1u : 0u;
        const u32 m2 = hw::ballot(wave::bor(v) != bor) != 0u ? 1u : 0u;
        const u32 m3 = static_cast<u32>(hw::popc(hw::ballot(wave::exclusiveSum(small) != bpre)))
            + static_cast<u32>(hw::popc(hw::ballot(wave::exclusiveSum(v) != vpre)));
        if(me == 0u)
        {
            if(m0)
                atom::add(&result[0], m0);
            if(m1)
                atom::add(&result[1], m1);
            if(m2)
                atom::add(&result[2], m2);
            if(m3)
                atom::add(&result[3], m3);
        }
    }
} // namespace

extern "C" {

int mmc_write_pattern(mmc_heap* heap, const uint64_t* d_offsets, const uint32_t* d_sizes, uint64_t n, uint64_t salt,
                      void* stream)
{
    if(!heap || (n && (!d_offsets || !d_sizes)))
        return mmc::fail(MMC_EINVAL, "null argument");
    if(n == 0)
        return MMC_OK;
    const unsigned threads = 256;
    u64 blocks = (n * 64u + threads - 1u) / threads;
    blocks = blocks < 16384u ? blocks : 16384u;
    hipLaunchKernelGGL(writePatternKernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                       static_cast<hipStream_t>(stream), heap->view, d_offsets, d_sizes, n, salt);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "write pattern launch");
}

int mmc_verify_pattern(mmc_heap* heap, const uint64_t* d_offsets, const uint32_t* d_sizes, uint64_t n,
                       uint64_t salt, uint64_t* d_result, void* stream)
{
    if(!heap || !d_result || (n && (!d_offsets || !d_sizes)))
        return mmc::fail(MMC_EINVAL, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(d_result, 0, 5 * sizeof(uint64_t), s);
    if(err != hipSuccess)
        return mmc::hipFail(err, "hipMemsetAsync");
    if(n == 0)
        return MMC_OK;
    const unsigned threads = 256; // four waves: the kernel's workgroup totals assume it
    u64 blocks = (n * 64u + threads - 1u) / threads;
    blocks = blocks < 2048u ? blocks : 2048u;
    hipLaunchKernelGGL(verifyPatternKernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0, s, heap->view,
                       d_offsets, d_sizes, n, salt, d_result);
    err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "verify pattern launch");
}

int mmc_alloc_until_null(mmc_heap* heap, uint32_t size, uint64_t* d_offsets, uint64_t cap, uint64_t* d_count,
                         uint32_t blocks, uint32_t threads, void* stream)
{
    if(!heap || !d_offsets || !d_count || blocks == 0 || threads == 0 || threads > 1024)
        return mmc::fail(MMC_EINVAL, "bad argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(d_count, 0, sizeof(uint64_t), s);
    if(err != hipSuccess)
        return mmc::hipFail(err, "hipMemsetAsync");
    hipLaunchKernelGGL(allocUntilNullKernel, dim3(blocks), dim3(threads), 0, s, heap->view, size, d_offsets, cap,
                       d_count);
    err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "alloc-until-null launch");
}

int mmc_bench_atomics(uint32_t* d_buf, uint64_t words, uint32_t pattern, uint32_t iters, uint32_t blocks,
                      uint32_t threads, uint64_t* d_sink, void* stream)
{
    if(!d_buf || !d_sink || words < 32 || blocks == 0 || threads == 0 || threads > 1024 || pattern > 3)
        return mmc::fail(MMC_EINVAL, "bad argument");
    hipLaunchKernelGGL(atomicsKernel, dim3(blocks), dim3(threads), 0, static_cast<hipStream_t>(stream), d_buf, words,
                       pattern, iters, d_sink);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "atomics launch");
}

int mmc_selftest_wave(uint32_t waves, uint64_t seed, uint32_t* d_result, void* stream)
{
    if(!d_result || waves == 0u)
        return mmc::fail(MMC_EINVAL, "bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t err = hipMemsetAsync(d_result, 0, 4u * sizeof(uint32_t), st);
    if(err != hipSuccess)
        return mmc::hipFail(err, "hipMemsetAsync(selftest)");
    hipLaunchKernelGGL(waveSelfTestKernel, dim3(waves), dim3(64), 0, st, seed, d_result);
    err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "selftest launch");
}

int mmc_selftest_fastdiv(uint64_t seed, uint64_t count, uint64_t* mismatches)
{
    if(!mismatches)
        return mmc::fail(MMC_EINVAL, "null output");
    uint64_t bad = 0;
    uint64_t z = seed;
    auto next = [&z]() {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    };
    auto check = [&bad](uint32_t n, uint32_t d) {
        const mallocMC::FastDiv f(d);
        if(f.div(n) != n / d || f.mod(n) != n % d)
            ++bad;
    };
    for(uint32_t k = 0; k < 32u; ++k)
        for(int64_t dd = -2; dd <= 2; ++dd)
        {
            const int64_t d64 = (int64_t{1} << k) + dd;
            if(d64 < 1 || d64 > 0xffffffffll)
                continue;
            const uint32_t d = static_cast<uint32_t>(d64);
            const uint32_t edges[] = {0u, 1u, d - 1u, d, d + 1u, 0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu,
                                      static_cast<uint32_t>((0xffffffffull / d) * d),
                                      static_cast<uint32_t>((0xffffffffull / d) * d - 1u)};
            for(uint32_t n : edges)
                check(n, d);
        }
    for(uint64_t i = 0; i < count; ++i)
    {
        const uint64_t r = next();
        // divisors spread over all magnitudes
        const uint32_t shift = static_cast<uint32_t>(r >> 59); // 0..31
        uint32_t d = static_cast<uint32_t>(r) >> shift;
        d = d == 0u ? 1u : d;
        const uint32_t n = static_cast<uint32_t>(next());
        check(n, d);
        check(static_cast<uint32_t>((static_cast<uint64_t>(n / d) * d)), d);
    }
    *mismatches = bad;
    return MMC_OK;
}

int mmc_bench_copy(const void* d_src, void* d_dst, uint64_t bytes, void* stream)
{
    if(!d_src || !d_dst || (bytes % 16u) != 0u)
        return mmc::fail(MMC_EINVAL, "bad argument");
    const u64 n16 = bytes / 16u;
    const unsigned threads = 256;
    u64 blocks = (n16 + threads - 1u) / threads;
    blocks = blocks < 8192u ? (blocks ? blocks : 1u) : 8192u;
    hipLaunchKernelGGL(copyKernel, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0,
                       static_cast<hipStream_t>(stream), static_cast<const uint4*>(d_src), static_cast<uint4*>(d_dst),
                       n16);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? MMC_OK : mmc::hipFail(err, "copy launch");
}

} // extern "C"
