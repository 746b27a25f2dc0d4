/*
 * AlignmentPolicies::Noop (src/include/mallocMC/alignmentPolicies/Noop.hpp:47-65).
 */
#pragma once

#include "../mallocMC_utils.hpp"

#include <cstddef>
#include <cstdint>
#include <string>
#include <tuple>

namespace mallocMC
{
    namespace AlignmentPolicies
    {
        class Noop
        {
        public:
            static constexpr std::size_t dataAlignment = 1;

            static std::tuple<void*, std::size_t> alignPool(void* memory, std::size_t memsize)
            {
                return std::make_tuple(memory, memsize);
            }

            MALLOCMC_HD static std::uint32_t applyPadding(std::uint32_t bytes)
            {
                return bytes;
            }

            static std::string classname()
            {
                return "Noop";
            }
        };
    } // namespace AlignmentPolicies
} // namespace mallocMC
