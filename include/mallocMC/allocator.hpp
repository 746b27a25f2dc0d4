/*
 * Host class of the heap (src/include/mallocMC/allocator.hpp:97-242):
 * Allocator<Acc, Creation, Distribution, OOM, ReservePool, Alignment>.
 * alpaka's device and queue become a HIP device ordinal and a hipStream_t.
 * Lifecycle as the reference's alloc() (:136-159): reserve the pool through
 * the ReservePool policy, align it with the Alignment policy, then let the
 * Creation policy build the heap in it (CreationPolicy::initHeap, which
 * enqueues the init kernel on the queue and fills the device allocator).
 * getAvailableSlots goes through CreationPolicy::getAvailableSlotsHost, as
 * the reference's does (:60-72).  Host errors throw, as the reference's do.
 */
#pragma once

#include "mallocmc_hip.h"

#include "device_allocator.hpp"
#include "mallocMC_allocator_handle.hpp"
#include "mallocMC_constraints.hpp"
#include "mallocMC_traits.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace mallocMC
{
    struct HeapInfo
    {
        void* p;
        std::size_t size;
    };

    template<
        typename T_Acc,
        typename T_CreationPolicy,
        typename T_DistributionPolicy,
        typename T_OOMPolicy,
        typename T_ReservePoolPolicy,
        typename T_AlignmentPolicy>
    class Allocator
        : public PolicyConstraints<
              T_CreationPolicy,
              T_DistributionPolicy,
              T_OOMPolicy,
              T_ReservePoolPolicy,
              T_AlignmentPolicy>
    {
    public:
        using CreationPolicy = T_CreationPolicy;
        using DistributionPolicy = T_DistributionPolicy;
        using OOMPolicy = T_OOMPolicy;
        using ReservePoolPolicy = T_ReservePoolPolicy;
        using AlignmentPolicy = T_AlignmentPolicy;
        using HeapInfoVector = std::vector<HeapInfo>;
        using DevAllocator = DeviceAllocator<CreationPolicy, DistributionPolicy, OOMPolicy, AlignmentPolicy>;
        using AllocatorHandle = AllocatorHandleImpl<Allocator>;

        Allocator(int device, hipStream_t queue, std::size_t size = 8u * 1024u * 1024u)
        {
            alloc(device, queue, size);
        }

        Allocator(const Allocator&) = delete;
        Allocator& operator=(const Allocator&) = delete;

        ~Allocator()
        {
            free();
        }

        //! destructiveResize (allocator.hpp:195-200)
        void destructiveResize(int device, hipStream_t queue, std::size_t size)
        {
            free();
            alloc(device, queue, size);
        }

        AllocatorHandle getAllocatorHandle() const
        {
            return AllocatorHandle{devAllocator};
        }

        operator AllocatorHandle() const
        {
            return getAllocatorHandle();
        }

        //! info (allocator.hpp:214-223)
        static std::string info(std::string linebreak = " ")
        {
            std::stringstream ss;
            ss << "CreationPolicy:      " << CreationPolicy::classname() << "    " << linebreak;
            ss << "DistributionPolicy:  " << DistributionPolicy::classname() << "" << linebreak;
            ss << "OOMPolicy:           " << OOMPolicy::classname() << "         " << linebreak;
            ss << "ReservePoolPolicy:   " << ReservePoolPolicy::classname() << " " << linebreak;
            ss << "AlignmentPolicy:     " << AlignmentPolicy::classname() << "   " << linebreak;
            return ss.str();
        }

        //! getAvailableSlots (allocator.hpp:227-233): pads, counts, blocks on the queue.
        unsigned getAvailableSlots(int device, hipStream_t queue, std::size_t slotSize)
        {
            slotSize = AlignmentPolicy::applyPadding(static_cast<std::uint32_t>(slotSize));
            if constexpr(!Traits<Allocator>::providesAvailableSlots)
                return 0u;
            else
                return CreationPolicy::template getAvailableSlotsHost<T_Acc>(device, queue, slotSize, heap);
        }

        //! getHeapLocations (allocator.hpp:235-241)
        HeapInfoVector getHeapLocations() const
        {
            HeapInfoVector v;
            v.push_back(heapInfos);
            return v;
        }

        //! The underlying C ABI heap, for the batch entry points of mallocmc_batched.h.
        mmc_heap* cHeap() const
        {
            return heap;
        }

    private:
        static void check(int rc, const char* what)
        {
            if(rc != MMC_OK)
            {
                const std::string msg = std::string(what) + ": " + mmc_last_error();
                if(rc == MMC_EINVAL)
                    throw std::invalid_argument(msg);
                throw std::runtime_error(msg);
            }
        }

        //! alloc (allocator.hpp:136-159)
        void alloc(int device, hipStream_t queue, std::size_t size)
        {
            if(hipSetDevice(device) != hipSuccess)
                throw std::runtime_error("hipSetDevice failed");
            void* const raw = reservePolicy.setMemPool(device, size);
            void* pool = raw;
            std::tie(pool, size) = AlignmentPolicy::alignPool(pool, size);
            try
            {
                heap = CreationPolicy::template initHeap<T_Acc>(device, queue, &devAllocator, pool, size);
            }
            catch(...)
            {
                reservePolicy.resetMemPool(raw);
                throw;
            }
            devAllocator.pool = pool;
            void* p = nullptr;
            std::uint64_t s = 0;
            check(mmc_heap_location(heap, &p, &s), "getHeapLocations");
            rawPool = raw; // what the reserve policy handed out, before alignment
            heapInfos.p = p;
            heapInfos.size = static_cast<std::size_t>(s);
        }

        //! free (allocator.hpp:166-172)
        void free()
        {
            if(heap)
                mmc_destroy(heap);
            heap = nullptr;
            if(rawPool)
                reservePolicy.resetMemPool(rawPool);
            rawPool = nullptr;
            heapInfos.p = nullptr;
            heapInfos.size = 0;
        }

        ReservePoolPolicy reservePolicy;
        mmc_heap* heap = nullptr;
        void* rawPool = nullptr;
        DevAllocator devAllocator{};
        HeapInfo heapInfos{nullptr, 0};
    };
} // namespace mallocMC
