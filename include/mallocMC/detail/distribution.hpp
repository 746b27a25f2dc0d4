/*
 * Distribution policies for wave64: the identity policy
 * (src/include/mallocMC/distributionPolicies/Noop.hpp:47-66) and XMallocSIMD
 * request coalescing (src/include/mallocMC/distributionPolicies/XMallocSIMD.hpp:118-176)
 * redone for 64-lane wavefronts.
 *
 * XMallocSIMD on wave64: the reference admits a lane when 0 < bytes <
 * pagesize/32, sized for 32 lanes; with 64 lanes the leader's request
 * (16 + sum) can exceed a page and become a page-based allocation whose
 * slices destroy() cannot attribute (XMallocSIMD.hpp:128-132 with
 * Scatter.hpp:924-940).  Here a lane is coalescible when 0 < bytes <
 * pagesize/64, so 16 + 64 * max(bytes) <= pagesize always holds.  Offsets are
 * assigned in lane order from an explicit ballot + readlane scan instead of
 * an LDS atomic that relies on lock-step execution (XMallocSIMD.hpp:141-147).
 */
#pragma once

#include "scatter_heap.hpp"

#include <type_traits>

namespace mallocMC
{
    namespace detail
    {
        //! Whether a distribution policy serves several lanes from one
        //! allocation (its `coalesces`; false for policies without one).
        template<class P, class = void>
        struct CoalescesOf : std::false_type
        {
        };
        template<class P>
        struct CoalescesOf<P, std::void_t<decltype(P::coalesces)>> : std::bool_constant<P::coalesces>
        {
        };
        template<class P>
        constexpr bool coalescesOf()
        {
            return CoalescesOf<P>::value;
        }

        //! Whether a distribution policy names the device type the
        //! wave-cooperative path runs (its `Device`); a user-written policy
        //! with only the concept members is served per lane.
        template<class P, class = void>
        struct HasWaveDevice : std::false_type
        {
        };
        template<class P>
        struct HasWaveDevice<P, std::void_t<typename P::Device>> : std::true_type
        {
        };

        struct NoopDistribution
        {
            template<class Heap>
            MALLOCMC_D u32 collect(const Heap&, u32 bytes)
            {
                return bytes;
            }
            MALLOCMC_D void* distribute(void* p)
            {
                return p;
            }
        };

        struct XMallocSIMD64
        {
            bool canCoalesce = false;
            u32 myoffset = 0;
            u32 threadcount = 0;
            u32 leader = 0;
            u32 reqSize = 0;

            //! collect (XMallocSIMD.hpp:118-150) with the heap's page bound;
            //! all lanes of the wave that will call distribute() must call
            //! collect() together.
            template<class Heap>
            MALLOCMC_D u32 collect(const Heap& heap, u32 bytes)
            {
                return collectBelow(heap.pagesize / waveSize, bytes);
            }

            //! collect with lanes of 0 < bytes < limit coalescible; limit =
            //! pagesize / 64 keeps 16 + 64 * max(bytes) within a page.
            MALLOCMC_D u32 collectBelow(u32 limit, u32 bytes)
            {
                const bool coalescible = bytes > 0u && bytes < limit;
                const u64 m = hw::ballot(coalescible);
                threadcount = hw::popc(m);
                canCoalesce = coalescible && threadcount > 1u;
                reqSize = bytes;
                if(threadcount > 1u)
                {
                    const u32 me = hw::laneid();
                    u32 prefix = 0u;
                    u32 total = 0u;
                    for(u64 mm = m; mm != 0u; mm &= mm - 1u)
                    {
                        const u32 l = hw::ffs(mm) - 1u;
                        const u32 b = __builtin_amdgcn_readlane(bytes, l);
                        prefix += l < me ? b : 0u;
                        total += b;
                    }
                    leader = hw::ffs(m) - 1u;
                    if(canCoalesce)
                    {
                        myoffset = 16u + prefix;
                        reqSize = me == leader ? 16u + total : 0u;
                    }
                }
                return reqSize;
            }

            //! distribute (XMallocSIMD.hpp:152-176)
            MALLOCMC_D void* distribute(void* allocated)
            {
                if(threadcount <= 1u)
                    return allocated;
                const u64 mine = reinterpret_cast<u64>(allocated);
                const u64 base = hw::readlane64(mine, leader);
                if(!canCoalesce)
                    return allocated;
                if(reqSize != 0u && allocated != nullptr)
                    atom::store(static_cast<u32*>(allocated), threadcount);
                return base == 0u ? nullptr : reinterpret_cast<void*>(base + myoffset);
            }
        };
    } // namespace detail
} // namespace mallocMC
