/*
 * Wave64 hardware helpers for gfx950 (CDNA4).
 *
 * Replaces the reference's hardware shim src/include/mallocMC/mallocMC_utils.hpp
 * (laneid :91-102, warpid :111-124, smid :126-137, lanemask_lt :139-150,
 * ballot :152-162, activemask :165-175, ffs :205-224, popc :226-245,
 * threadfence :247-293).  Lane masks are 64-bit: a gfx950 wavefront has 64
 * lanes.  All shared heap state is accessed through agent-scope atomics so
 * that the eight XCDs' L2 caches never serve a stale copy.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define MALLOCMC_HD __host__ __device__ __forceinline__
#define MALLOCMC_D __device__ __forceinline__

namespace mallocMC
{
    //! Wavefront width of gfx950.
    constexpr std::uint32_t waveSize = 64u;
    //! Largest work-group the HIP runtime launches (used to size per-wave LDS arrays).
    constexpr std::uint32_t maxThreadsPerBlock = 1024u;

    //! Accelerator tag of the gfx950 device: alpaka's accelerator object has
    //! no counterpart in HIP device code, so the `acc` the policies take is a
    //! tag; device functions accept any type.
    struct AccGfx950
    {
    };

    template<class T>
    MALLOCMC_HD T divup(T a, T b)
    {
        return (a + b - 1) / b;
    }

    //! Unsigned 32-bit division by a divisor fixed at heap set-up, as a
    //! multiply-high and shifts (Granlund & Montgomery, "Division by
    //! invariant integers using multiplication", PLDI 1994, the round-up
    //! variant with the added bit): exact for every 32-bit dividend.
    struct FastDiv
    {
        std::uint32_t d = 1u;
        std::uint32_t m = 0u; // 0: d is a power of two, q = n >> s
        std::uint32_t s = 0u;

        MALLOCMC_HD FastDiv() = default;

        MALLOCMC_HD explicit FastDiv(std::uint32_t divisor)
        {
            d = divisor == 0u ? 1u : divisor;
            if((d & (d - 1u)) == 0u)
            {
                m = 0u;
                s = 0u;
                while((1u << s) < d)
                    ++s;
            }
            else
            {
                s = 0u;
                while((std::uint64_t{1} << s) < d)
                    ++s; // 2^(s-1) < d < 2^s
                const std::uint64_t num = (std::uint64_t{1} << 32) * ((std::uint64_t{1} << s) - d);
                m = static_cast<std::uint32_t>(num / d + 1u);
            }
        }

        MALLOCMC_HD std::uint32_t div(std::uint32_t n) const
        {
            if(m == 0u)
                return n >> s;
            const std::uint32_t t = static_cast<std::uint32_t>((static_cast<std::uint64_t>(n) * m) >> 32);
            return (t + ((n - t) >> 1)) >> (s - 1u);
        }

        MALLOCMC_HD std::uint32_t mod(std::uint32_t n) const
        {
            return n - div(n) * d;
        }
    };

    //! n / d for n < 2^24 and d >= 1 with a float reciprocal and one
    //! correction each way.  Exact whenever the quotient is below 2^22 (the
    //! reciprocal and the product together err by less than 2^-22 relative,
    //! so the first quotient is within one of the exact one); larger
    //! quotients may be off by a few units, which callers that clamp the
    //! result to 32 or less never see.
    MALLOCMC_D std::uint32_t udiv24(std::uint32_t n, std::uint32_t d)
    {
        std::uint32_t q = static_cast<std::uint32_t>(static_cast<float>(n) * __builtin_amdgcn_rcpf(static_cast<float>(d)));
        int r = static_cast<int>(n) - static_cast<int>(q * d);
        if(r < 0)
        {
            --q;
            r += static_cast<int>(d);
        }
        if(r >= static_cast<int>(d))
            ++q;
        return q;
    }

    namespace hw
    {
        MALLOCMC_D std::uint32_t laneid()
        {
            return __lane_id();
        }

        //! Wave slot within its SIMD (HW_ID[3:0]); changes only between launches.
        MALLOCMC_D std::uint32_t waveslot()
        {
            return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 4);
        }

        //! Wave slot and SIMD within the CU (HW_ID[5:0]): unlike waveslot(),
        //! distinct for waves of one CU that sit in the same slot of
        //! different SIMDs.
        MALLOCMC_D std::uint32_t waveInCu()
        {
            return __builtin_amdgcn_s_getreg((5 << 11) | (0 << 6) | 4);
        }

        //! Compute unit id including SE and XCD bits (HIP's __smid on gfx950).
        MALLOCMC_D std::uint32_t cuid()
        {
            return __smid();
        }

        //! XCD (XCC) the wave runs on, 0..7 on MI355X.
        MALLOCMC_D std::uint32_t xccid()
        {
            return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);
        }

        MALLOCMC_D std::uint64_t ballot(bool pred)
        {
            return __ballot(pred);
        }

        MALLOCMC_D std::uint64_t activemask()
        {
            return __ballot(1);
        }

        MALLOCMC_D std::uint64_t lanemask_lt()
        {
            return __lanemask_lt();
        }

        //! 1-based index of the lowest set bit, 0 for 0 (the reference's ffs contract).
        MALLOCMC_D std::uint32_t ffs(std::uint64_t m)
        {
            return static_cast<std::uint32_t>(__ffsll(static_cast<unsigned long long>(m)));
        }

        MALLOCMC_D std::uint32_t popc(std::uint64_t m)
        {
            return static_cast<std::uint32_t>(__popcll(static_cast<unsigned long long>(m)));
        }

        //! Set bits of m below this lane, popc(m & ((1 << lane) - 1)): two
        //! mbcnt instructions.  Equals popc(m & lanemask_lt()) when m holds
        //! bits of active lanes only (a ballot or a subset of one).
        MALLOCMC_D std::uint32_t rankBelow(std::uint64_t m)
        {
            return __builtin_amdgcn_mbcnt_hi(
                static_cast<std::uint32_t>(m >> 32),
                __builtin_amdgcn_mbcnt_lo(static_cast<std::uint32_t>(m), 0u));
        }

        template<class T>
        MALLOCMC_D T readlane(T v, std::uint32_t lane)
        {
            return __shfl(v, static_cast<int>(lane), 64);
        }

        //! 64-bit value of lane `lane`; `lane` must be the same in all lanes.
        MALLOCMC_D std::uint64_t readlane64(std::uint64_t v, std::uint32_t lane)
        {
            const std::uint32_t lo = static_cast<std::uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(static_cast<std::uint32_t>(v)), static_cast<int>(lane)));
            const std::uint32_t hi = static_cast<std::uint32_t>(
                __builtin_amdgcn_readlane(static_cast<int>(static_cast<std::uint32_t>(v >> 32)), static_cast<int>(lane)));
            return (static_cast<std::uint64_t>(hi) << 32) | lo;
        }

        //! Makes the wave wait for `v` (the result of returning atomics) before
        //! any later memory operation is issued: the empty asm reads the
        //! register, so the compiler must wait for the data, and its memory
        //! clobber keeps later atomics from being moved above it.  A returning
        //! device-scope atomic has been performed at the device's coherence
        //! point once its result is back, so every atomic issued after this
        //! point is ordered after it for all observers that use atomics on
        //! those words.  No L2 write-back is involved (an agent-scope release
        //! on gfx950 writes back the whole L2 of the issuing XCD).
        MALLOCMC_D void consume(std::uint32_t v)
        {
            asm volatile("" ::"v"(v) : "memory");
        }
    } // namespace hw

    //! Agent-scope relaxed atomics on heap metadata (the reference uses
    //! alpaka::atomicOp with the default grid hierarchy, alpaka/atomic/Traits.hpp:70-78).
    namespace atom
    {
        template<class T>
        MALLOCMC_D T load(const T* p)
        {
            return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D void store(T* p, T v)
        {
            __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T add(T* p, T v)
        {
            return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T sub(T* p, T v)
        {
            return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T or_(T* p, T v)
        {
            return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T and_(T* p, T v)
        {
            return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T exch(T* p, T v)
        {
            return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T min(T* p, T v)
        {
            return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        template<class T>
        MALLOCMC_D T max(T* p, T v)
        {
            return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        //! Release-ordered subtract: every store of this lane before it is
        //! visible to any agent lane that observes its result.
        template<class T>
        MALLOCMC_D T sub_release(T* p, T v)
        {
            return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        //! Compare-and-swap returning the old value (CUDA atomicCAS contract).
        template<class T>
        MALLOCMC_D T cas(T* p, T cmp, T val)
        {
            __hip_atomic_compare_exchange_strong(
                p,
                &cmp,
                val,
                __ATOMIC_RELAXED,
                __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
            return cmp;
        }
        //! Agent-scope fence (the reference's threadfenceDevice, mallocMC_utils.hpp:247-293).
        MALLOCMC_D void fence_device()
        {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        }
    } // namespace atom
} // namespace mallocMC
