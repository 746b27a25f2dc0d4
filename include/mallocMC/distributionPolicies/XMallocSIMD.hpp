/*
 * DistributionPolicies::XMallocSIMD
 * (src/include/mallocMC/distributionPolicies/XMallocSIMD.hpp:49-184): the small
 * requests of a wavefront are served by one allocation with a 16-byte header
 * counting the slices.  On wave64 a lane joins when 0 < bytes < pagesize/64
 * (detail::XMallocSIMD64 explains why the reference's pagesize/32 bound
 * overflows a page with 64 lanes).
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
            using Device = detail::XMallocSIMD64;
            static constexpr bool coalesces = true;
#ifdef MALLOCMC_DP_XMALLOCSIMD_PAGESIZE
            static constexpr std::uint32_t pagesize = MALLOCMC_DP_XMALLOCSIMD_PAGESIZE;
#else
            static constexpr std::uint32_t pagesize = Properties::pagesize;
#endif
            static constexpr std::uint32_t _pagesize = pagesize;

            static std::string classname()
            {
                std::stringstream ss;
                ss << "XMallocSIMD[" << pagesize << "]";
                return ss.str();
            }
        };
    } // namespace DistributionPolicies
} // namespace mallocMC
