/*
 * DeviceAllocator: the policies composed into device-side malloc / free /
 * getAvailableSlots (src/include/mallocMC/device_allocator.hpp:55-107):
 * pad -> collect -> create -> isOOM -> handleOOM -> distribute.
 *
 * The object is trivially copyable and carries the heap view by value, so a
 * kernel receives it as an argument (no pointer chase to a device-resident
 * allocator object).  `acc` parameters are accepted for source compatibility
 * and ignored: HIP device code has no accelerator object.
 */
#pragma once

#include "detail/device_allocator.hpp"
#include "detail/wave_alloc.hpp"

#include <cstddef>
#include <cstdint>

namespace mallocMC
{
    template<typename T_CreationPolicy, typename T_DistributionPolicy, typename T_OOMPolicy, typename T_AlignmentPolicy>
    class DeviceAllocator
    {
    public:
        using CreationPolicy = T_CreationPolicy;
        using DistributionPolicy = T_DistributionPolicy;
        using OOMPolicy = T_OOMPolicy;
        using AlignmentPolicy = T_AlignmentPolicy;

        detail::ScatterHeap heap;
        void* pool = nullptr;

        //! DeviceAllocator::malloc (device_allocator.hpp:72-82).  Callable from
        //! any lane; with XMallocSIMD the lanes of a wavefront that call it
        //! together share one allocation.  When all 64 lanes of the wavefront
        //! call it together (full active mask) and none asks for more than a
        //! page, they are served by the wave-cooperative path (malloc_wave):
        //! same heap protocol, one group claim per size class instead of one
        //! reference walk per lane.  Otherwise each lane runs the reference's
        //! per-lane algorithm (Scatter.hpp:895-912 and what it calls); its
        //! page-based path serializes runs of pages on one mutex and keeps
        //! _firstFreePageBased exact, so runs stay one short scan each.
        template<typename Acc>
        __device__ void* malloc(const Acc&, std::size_t bytes) const
        {
            return malloc(bytes);
        }

        __device__ void* malloc(std::size_t bytes) const
        {
            if(hw::activemask() == ~std::uint64_t{0} && hw::ballot(bytes > heap.pagesize) == 0u)
                return malloc_wave(bytes);
            return mallocLane(bytes);
        }

        //! The reference's per-lane malloc, whatever the active mask.
        __device__ void* mallocLane(std::size_t bytes) const
        {
            const std::uint32_t padded = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(bytes));
            typename DistributionPolicy::Device dist;
            const std::uint32_t req = dist.collect(heap, padded);
            void* mem = heap.create<detail::WaveIds>(req);
            if(detail::isOOM(mem, req))
                mem = OOMPolicy::handleOOM(mem);
            return dist.distribute(mem);
        }

        //! DeviceAllocator::free (device_allocator.hpp:84-88).
        template<typename Acc>
        __device__ void free(const Acc&, void* p) const
        {
            free(p);
        }

        //! With a full active mask the wavefront frees cooperatively
        //! (free_wave: frees grouped by page); otherwise per lane.
        __device__ void free(void* p) const
        {
            if(hw::activemask() == ~std::uint64_t{0})
                free_wave(p);
            else
                freeLane(p);
        }

        //! The reference's per-lane free (Scatter.hpp:918-946).
        __device__ void freeLane(void* p) const
        {
            heap.destroy<detail::WaveIds>(p);
        }

        /** Wave-cooperative malloc (no counterpart in the reference): all 64
         * lanes of the wavefront must call it together in uniform control
         * flow; lanes without a request pass 0.  Same heap protocol as
         * malloc(), so either free() or free_wave() releases the result. */
        __device__ void* malloc_wave(std::size_t bytes, detail::WaveTuning tuning = {}) const
        {
            const std::uint32_t padded = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(bytes));
            const detail::WaveAlloc wa{heap, tuning};
            // the wave path pads with the heap's alignment, equal to AlignmentPolicy's
            void* mem = wa.malloc<typename DistributionPolicy::Device>(padded);
            if(detail::isOOM(mem, padded))
                mem = OOMPolicy::handleOOM(mem);
            return mem;
        }

        //! Wave-cooperative free; all 64 lanes call it together (nullptr allowed).
        __device__ void free_wave(void* p, detail::WaveTuning tuning = {}) const
        {
            const detail::WaveAlloc wa{heap, tuning};
            wa.free(p);
        }

        /** getAvailableSlots from device code (device_allocator.hpp:98-106 ->
         * Scatter.hpp:1311-1352): the active lanes of the calling wavefront
         * scan the heap together and each receives the total.  Every active
         * lane of the wave must call it (the reference asks the same of all
         * threads of a block). */
        template<typename Acc>
        __device__ unsigned getAvailableSlots(const Acc&, std::size_t slotSize) const
        {
            return getAvailableSlots(slotSize);
        }

        __device__ unsigned getAvailableSlots(std::size_t slotSize) const
        {
            const std::uint32_t s = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(slotSize));
            if(s == 0u)
                return 0u;
            const std::uint64_t active = hw::activemask();
            const std::uint32_t n = hw::popc(active);
            const std::uint32_t id = hw::popc(active & hw::lanemask_lt());
            const std::uint32_t mine = heap.availableSlotsPartial(s, id, n);
            std::uint32_t total = 0u;
            for(std::uint64_t m = active; m != 0u; m &= m - 1u)
                total += __builtin_amdgcn_readlane(mine, hw::ffs(m) - 1u);
            return total;
        }
    };
} // namespace mallocMC
