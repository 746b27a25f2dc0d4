/*
 * DistributionPolicies::XMallocSIMD
 * (src/include/mallocMC/distributionPolicies/XMallocSIMD.hpp:49-184): the small
 * requests of a wavefront are served by one allocation with a 16-byte header
 * counting the slices.  On wave64 a lane joins when 0 < bytes < pagesize/64
 * (detail::XMallocSIMD64 explains why the reference's pagesize/32 bound
 * overflows a page with 64 lanes).  collect / distribute take the
 * accelerator, as the reference's members do (:90-176).
 */
#pragma once

#include "../detail/distribution.hpp"

#include <cstdint>
#include <sstream>
#include <string>

namespace mallocMC
{
    namespace DistributionPolicies
    {
        namespace XMallocSIMDConf
        {
            struct DefaultXMallocConfig
            {
                static constexpr auto pagesize = 4096;
            };
        } // namespace XMallocSIMDConf

        template<typename T_Config = XMallocSIMDConf::DefaultXMallocConfig>
        class XMallocSIMD
        {
        public:
            using Properties = T_Config;
            //! the same policy for the wave-cooperative path
            using Device = detail::XMallocSIMD64;
            static constexpr bool coalesces = true;
#ifdef MALLOCMC_DP_XMALLOCSIMD_PAGESIZE
            static constexpr std::uint32_t pagesize = MALLOCMC_DP_XMALLOCSIMD_PAGESIZE;
#else
            static constexpr std::uint32_t pagesize = Properties::pagesize;
#endif
            static constexpr std::uint32_t _pagesize = pagesize;

            MALLOCMC_HD XMallocSIMD() = default;

            //! XMallocSIMD(acc) (XMallocSIMD.hpp:90-98): one object per call.
            template<typename Acc>
            MALLOCMC_D explicit XMallocSIMD(const Acc&)
            {
            }

            //! collect (XMallocSIMD.hpp:118-150): a lane joins when
            //! 0 < bytes < pagesize / 64; the first joining lane requests 16 +
            //! the sum of the joining lanes' sizes, the others 0.  All lanes
            //! of the wavefront that call distribute() call collect() together.
            template<typename Acc>
            MALLOCMC_D std::uint32_t collect(const Acc&, std::uint32_t bytes)
            {
                return state.collectBelow(pagesize / waveSize, bytes);
            }

            //! distribute (XMallocSIMD.hpp:152-176): the leader writes the
            //! slice count into the 16-byte header; every joining lane gets
            //! chunk + 16 + the sizes of the joining lanes below it.
            template<typename Acc>
            MALLOCMC_D void* distribute(const Acc&, void* allocatedMem)
            {
                return state.distribute(allocatedMem);
            }

            static std::string classname()
            {
                std::stringstream ss;
                ss << "XMallocSIMD[" << pagesize << "]";
                return ss.str();
            }

        private:
            detail::XMallocSIMD64 state;
        };
    } // namespace DistributionPolicies
} // namespace mallocMC
