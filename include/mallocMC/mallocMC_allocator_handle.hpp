/*
 * AllocatorHandleImpl: what a kernel receives
 * (src/include/mallocMC/mallocMC_allocator_handle.hpp:35-63).  Here it holds
 * the device allocator by value instead of a pointer to a device-resident
 * copy; it stays valid until the Allocator is destroyed or resized.
 */
#pragma once

#include <cstddef>

namespace mallocMC
{
    template<typename T_HostAllocator>
    struct AllocatorHandleImpl
    {
        using DevAllocator = typename T_HostAllocator::DevAllocator;

        DevAllocator devAllocator;

        explicit AllocatorHandleImpl(const DevAllocator& d) : devAllocator(d)
        {
        }

        template<typename Acc>
        __device__ void* malloc(const Acc& acc, std::size_t size) const
        {
            return devAllocator.malloc(acc, size);
        }

        template<typename Acc>
        __device__ void free(const Acc& acc, void* p) const
        {
            devAllocator.free(acc, p);
        }

        template<typename Acc>
        __device__ unsigned getAvailableSlots(const Acc& acc, std::size_t slotSize) const
        {
            return devAllocator.getAvailableSlots(acc, slotSize);
        }

        __device__ void* malloc(std::size_t size) const
        {
            return devAllocator.malloc(size);
        }

        __device__ void free(void* p) const
        {
            devAllocator.free(p);
        }

        __device__ unsigned getAvailableSlots(std::size_t slotSize) const
        {
            return devAllocator.getAvailableSlots(slotSize);
        }
    };
} // namespace mallocMC
