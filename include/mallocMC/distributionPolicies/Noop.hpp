/*
 * DistributionPolicies::Noop (src/include/mallocMC/distributionPolicies/Noop.hpp:47-66):
 * every lane allocates for itself.
 */
#pragma once

#include "../detail/distribution.hpp"

#include <string>

namespace mallocMC
{
    namespace DistributionPolicies
    {
        class Noop
        {
        public:
            //! device implementation used by DeviceAllocator
            using Device = detail::NoopDistribution;
            static constexpr bool coalesces = false;

            static std::string classname()
            {
                return "Noop";
            }
        };
    } // namespace DistributionPolicies
} // namespace mallocMC
