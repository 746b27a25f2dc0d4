/*
 * Composition of the policies into device-side malloc / free /
 * getAvailableSlots, as DeviceAllocator does in
 * src/include/mallocMC/device_allocator.hpp:72-106: pad -> collect -> create
 * -> isOOM -> handleOOM -> distribute.
 */
#pragma once

#include "distribution.hpp"
#include "scatter_heap.hpp"

namespace mallocMC
{
    namespace detail
    {
        //! ReturnNull::handleOOM (src/include/mallocMC/oOMPolicies/ReturnNull.hpp:47-51)
        struct ReturnNullOOM
        {
            static MALLOCMC_D void* handleOOM(void*)
            {
                return nullptr;
            }
        };

        //! Scatter::isOOM (Scatter.hpp:1051-1055)
        MALLOCMC_D bool isOOM(const void* p, u32 s)
        {
            return s != 0u && p == nullptr;
        }

        //! DeviceAllocator::malloc (device_allocator.hpp:72-82).
        template<class Ids, class Dist, class OOM = ReturnNullOOM>
        MALLOCMC_D void* deviceMalloc(const ScatterHeap& heap, u64 bytes)
        {
            const u32 padded = heap.pad(static_cast<u32>(bytes));
            Dist dist;
            const u32 req = dist.collect(heap, padded);
            void* mem = heap.create<Ids>(req);
            if(isOOM(mem, req))
                mem = OOM::handleOOM(mem);
            return dist.distribute(mem);
        }

        //! DeviceAllocator::free (device_allocator.hpp:84-88).
        template<class Ids>
        MALLOCMC_D void deviceFree(const ScatterHeap& heap, void* p)
        {
            heap.destroy<Ids>(p);
        }
    } // namespace detail
} // namespace mallocMC
