/*
 * DistributionPolicies::Noop (src/include/mallocMC/distributionPolicies/Noop.hpp:47-72):
 * every lane allocates for itself.  The members follow the reference's
 * distribution-policy concept -- constructed from the accelerator per call,
 * collect(acc, bytes) before the creation policy's create, distribute(acc,
 * mem) after it -- so a user-written policy with the same members plugs into
 * DeviceAllocator in its place.
 */
#pragma once

#include "../detail/distribution.hpp"

#include <cstdint>
#include <string>

namespace mallocMC
{
    namespace DistributionPolicies
    {
        class Noop
        {
        public:
            //! the same policy for the wave-cooperative path
            using Device = detail::NoopDistribution;
            static constexpr bool coalesces = false;

            MALLOCMC_HD Noop() = default;

            template<typename Acc>
            MALLOCMC_D explicit Noop(const Acc&)
            {
            }

            template<typename Acc>
            MALLOCMC_D std::uint32_t collect(const Acc&, std::uint32_t bytes) const
            {
                return bytes;
            }

            template<typename Acc>
            MALLOCMC_D void* distribute(const Acc&, void* allocatedMem) const
            {
                return allocatedMem;
            }

            static std::string classname()
            {
                return "Noop";
            }
        };
    } // namespace DistributionPolicies
} // namespace mallocMC
