/*
 * AlignmentPolicies::Shrink (src/include/mallocMC/alignmentPolicies/Shrink.hpp:60-145):
 * the pool start is rounded up to dataAlignment (the size shrinks by
 * dataAlignment + misalignment) and requests are padded to a multiple of it.
 */
#pragma once

#include "../mallocMC_utils.hpp"

#include <cstddef>
#include <cstdint>
#include <iostream>
#include <sstream>
#include <string>
#include <tuple>

namespace mallocMC
{
    namespace AlignmentPolicies
    {
        namespace ShrinkConfig
        {
            struct DefaultShrinkConfig
            {
                static constexpr auto dataAlignment = 16;
            };
        } // namespace ShrinkConfig

        template<typename T_Config = ShrinkConfig::DefaultShrinkConfig>
        class Shrink
        {
        public:
            using Properties = T_Config;
#ifdef MALLOCMC_AP_SHRINK_DATAALIGNMENT
            static constexpr std::size_t dataAlignment = MALLOCMC_AP_SHRINK_DATAALIGNMENT;
#else
            static constexpr std::size_t dataAlignment = Properties::dataAlignment;
#endif
            static_assert(
                dataAlignment != 0 && (dataAlignment & (dataAlignment - 1)) == 0,
                "dataAlignment must also be a power of 2");

            static std::tuple<void*, std::size_t> alignPool(void* memory, std::size_t memsize)
            {
                const auto status = reinterpret_cast<std::uintptr_t>(memory) & (dataAlignment - 1);
                if(status != 0)
                {
                    std::cout << "Heap Warning: memory to use not " << dataAlignment << " byte aligned...\n";
                    memory = reinterpret_cast<void*>(reinterpret_cast<std::uintptr_t>(memory) + dataAlignment - status);
                    memsize -= dataAlignment + static_cast<std::size_t>(status);
                }
                return std::make_tuple(memory, memsize);
            }

            MALLOCMC_HD static std::uint32_t applyPadding(std::uint32_t bytes)
            {
                constexpr std::uint32_t bitsToClear = static_cast<std::uint32_t>(dataAlignment) - 1u;
                return (bytes + bitsToClear) & ~bitsToClear;
            }

            static std::string classname()
            {
                std::stringstream ss;
                ss << "Shrink[" << dataAlignment << "]";
                return ss.str();
            }
        };
    } // namespace AlignmentPolicies
} // namespace mallocMC
