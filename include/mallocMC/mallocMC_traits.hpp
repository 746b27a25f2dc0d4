/*
 * Traits (src/include/mallocMC/mallocMC_traits.hpp:33-37).
 */
#pragma once

namespace mallocMC
{
    template<class T_Allocator>
    struct Traits
    {
        static constexpr bool providesAvailableSlots = T_Allocator::CreationPolicy::providesAvailableSlots;
    };
} // namespace mallocMC
