/*
 * DeviceAllocator: the policies composed into device-side malloc / free /
 * getAvailableSlots (src/include/mallocMC/device_allocator.hpp:55-107):
 * pad -> collect -> create -> isOOM -> handleOOM -> distribute.
 *
 * As in the reference, DeviceAllocator derives from its creation policy and
 * calls the policies only through their members:
 *   CreationPolicy      create<AlignmentPolicy>(acc, bytes), destroy(acc, p),
 *                       static isOOM(p, bytes),
 *                       getAvailableSlotsAccelerator<AlignmentPolicy>(acc, n)
 *                       and providesAvailableSlots;
 *   DistributionPolicy  constructed from acc once per call, then
 *                       collect(acc, bytes) and distribute(acc, mem);
 *   OOMPolicy           static handleOOM(mem);
 *   AlignmentPolicy     static applyPadding(bytes).
 * A user-written distribution or OOM policy with these members plugs in.
 *
 * On top of the reference's per-lane algorithm, a wavefront whose 64 lanes
 * call malloc / free together is served cooperatively (malloc_wave /
 * free_wave) when the distribution policy names the device type of that path
 * (`Device`, as Noop and XMallocSIMD do); policies without one are always
 * served per lane.
 *
 * The object is trivially copyable and carries the heap view by value, so a
 * kernel receives it as an argument (no pointer chase to a device-resident
 * allocator object).  The accelerator argument is a tag: HIP device code
 * needs no accelerator object, and the calls without one pass AccGfx950{}.
 */
#pragma once

#include "detail/device_allocator.hpp"
#include "detail/wave_alloc.hpp"
#include "mallocMC_traits.hpp"

#include <cstddef>
#include <cstdint>

namespace mallocMC
{
    template<typename T_CreationPolicy, typename T_DistributionPolicy, typename T_OOMPolicy, typename T_AlignmentPolicy>
    class DeviceAllocator : public T_CreationPolicy
    {
    public:
        using CreationPolicy = T_CreationPolicy;
        using DistributionPolicy = T_DistributionPolicy;
        using OOMPolicy = T_OOMPolicy;
        using AlignmentPolicy = T_AlignmentPolicy;

        //! whether a full wavefront is served by the wave-cooperative path
        static constexpr bool waveCapable = detail::HasWaveDevice<DistributionPolicy>::value;

        void* pool = nullptr;

        /** DeviceAllocator::malloc (device_allocator.hpp:72-82).  Callable
         * from any lane.  When all 64 lanes of the wavefront call it together
         * (full active mask), none asks for more than a page and the
         * distribution policy has a wave device type, they are served by the
         * wave-cooperative path (malloc_wave): same heap protocol, one group
         * claim per size class instead of one reference walk per lane.
         * Otherwise each lane runs the reference's composition (mallocLane).
         * Requests above a page stay on the reference's page-based path: in
         * a kernel where some wavefronts are full and others are not (lanes
         * that leave an allocate-until-null loop at their own null), the
         * per-lane page-by-page scan needs the page-based cursor exact --
         * every page above it in use -- which concurrent wave run claims
         * cannot keep; with wave runs for the full wavefronts, 128 waves
         * filling 1 GiB with 2-page runs took over 150 s. */
        template<typename Acc>
        __device__ void* malloc(const Acc& acc, std::size_t bytes) const
        {
            if constexpr(waveCapable)
            {
                if(hw::activemask() == ~std::uint64_t{0} && hw::ballot(bytes > this->heap.pagesize) == 0u)
                    return malloc_wave(bytes);
            }
            return mallocLane(acc, bytes);
        }

        __device__ void* malloc(std::size_t bytes) const
        {
            return malloc(AccGfx950{}, bytes);
        }

        //! The reference's per-lane composition, whatever the active mask
        //! (device_allocator.hpp:72-82 line for line).
        template<typename Acc>
        __device__ void* mallocLane(const Acc& acc, std::size_t bytes) const
        {
            const std::uint32_t padded = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(bytes));
            DistributionPolicy distributionPolicy(acc);
            const std::uint32_t reqSize = distributionPolicy.collect(acc, padded);
            void* memBlock = CreationPolicy::template create<AlignmentPolicy>(acc, reqSize);
            if(CreationPolicy::isOOM(memBlock, reqSize))
                memBlock = OOMPolicy::handleOOM(memBlock);
            return distributionPolicy.distribute(acc, memBlock);
        }

        __device__ void* mallocLane(std::size_t bytes) const
        {
            return mallocLane(AccGfx950{}, bytes);
        }

        //! DeviceAllocator::free (device_allocator.hpp:84-88).  With a full
        //! active mask (and a wave-capable distribution policy) the wavefront
        //! frees cooperatively (free_wave: frees grouped by page).
        template<typename Acc>
        __device__ void free(const Acc& acc, void* p) const
        {
            if constexpr(waveCapable)
            {
                if(hw::activemask() == ~std::uint64_t{0})
                {
                    free_wave(p);
                    return;
                }
            }
            freeLane(acc, p);
        }

        __device__ void free(void* p) const
        {
            free(AccGfx950{}, p);
        }

        //! The reference's per-lane free: the creation policy's destroy.
        template<typename Acc>
        __device__ void freeLane(const Acc& acc, void* p) const
        {
            CreationPolicy::destroy(acc, p);
        }

        __device__ void freeLane(void* p) const
        {
            freeLane(AccGfx950{}, p);
        }

        /** Wave-cooperative malloc (no counterpart in the reference): all 64
         * lanes of the wavefront must call it together in uniform control
         * flow; lanes without a request pass 0.  Same heap protocol as
         * malloc(), so either free() or free_wave() releases the result.
         * Requires a distribution policy with a wave device type. */
        __device__ void* malloc_wave(std::size_t bytes, detail::WaveTuning tuning = {}) const
        {
            static_assert(waveCapable, "the wave path needs a distribution policy with a Device type");
            const std::uint32_t padded = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(bytes));
            const detail::WaveAlloc wa{this->heap, tuning};
            // the wave path pads with the heap's alignment, equal to AlignmentPolicy's
            void* mem = wa.template malloc<typename DistributionPolicy::Device>(padded);
            if(CreationPolicy::isOOM(mem, padded))
                mem = OOMPolicy::handleOOM(mem);
            return mem;
        }

        //! Wave-cooperative free; all 64 lanes call it together (nullptr allowed).
        __device__ void free_wave(void* p, detail::WaveTuning tuning = {}) const
        {
            const detail::WaveAlloc wa{this->heap, tuning};
            wa.free(p);
        }

        /** getAvailableSlots from device code (device_allocator.hpp:98-106):
         * pads the slot size and asks the creation policy, or returns 0 when
         * it provides no count. */
        template<typename Acc>
        __device__ unsigned getAvailableSlots(const Acc& acc, std::size_t slotSize) const
        {
            slotSize = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(slotSize));
            if constexpr(Traits<DeviceAllocator>::providesAvailableSlots)
                return CreationPolicy::template getAvailableSlotsAccelerator<AlignmentPolicy>(acc, slotSize);
            else
                return 0u;
        }

        __device__ unsigned getAvailableSlots(std::size_t slotSize) const
        {
            return getAvailableSlots(AccGfx950{}, slotSize);
        }
    };
} // namespace mallocMC
