/*
 * CreationPolicies::Scatter: the ScatterAlloc heap policy with the
 * reference's configuration surface (src/include/mallocMC/creationPolicies/Scatter.hpp:56-207):
 * compile-time config structs, overridable by -D MALLOCMC_CP_SCATTER_* with
 * the precedence default struct < template struct < command-line macro.  The
 * values become the runtime geometry of detail::ScatterHeap.
 *
 * The policy holds the heap's device view and has the reference's members:
 * create / destroy / isOOM / getAvailableSlotsAccelerator on the device
 * (Scatter.hpp:895-946, 1051-1055, 1311-1352), initHeap /
 * getAvailableSlotsHost / classname on the host (:1057-1150, 1235-1306,
 * 1354-1370).  DeviceAllocator derives from it, as the reference's does
 * (device_allocator.hpp:55-56), and composes these members with the
 * distribution, OOM and alignment policies.
 */
#pragma once

#include "../detail/scatter_heap.hpp"
#include "mallocmc_hip.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

namespace mallocMC
{
    namespace CreationPolicies
    {
        namespace ScatterConf
        {
            struct DefaultScatterConfig
            {
                static constexpr auto pagesize = 4096;
                static constexpr auto accessblocksize = 2u * 1024u * 1024u * 1024u;
                static constexpr auto regionsize = 16;
                static constexpr auto wastefactor = 2;
                static constexpr auto resetfreedpages = false;
            };

            struct DefaultScatterHashingParams
            {
                static constexpr auto hashingK = 38183;
                static constexpr auto hashingDistMP = 17497;
                static constexpr auto hashingDistWP = 1;
                static constexpr auto hashingDistWPRel = 1;
            };
        } // namespace ScatterConf

        template<
            class T_Config = ScatterConf::DefaultScatterConfig,
            class T_Hashing = ScatterConf::DefaultScatterHashingParams>
        class Scatter
        {
        public:
            using HeapProperties = T_Config;
            using HashingProperties = T_Hashing;
            struct Properties
                : HeapProperties
                , HashingProperties
            {
            };
            static constexpr bool providesAvailableSlots = true;

#ifdef MALLOCMC_CP_SCATTER_PAGESIZE
            static constexpr std::uint32_t pagesize = MALLOCMC_CP_SCATTER_PAGESIZE;
#else
            static constexpr std::uint32_t pagesize = HeapProperties::pagesize;
#endif
#ifdef MALLOCMC_CP_SCATTER_ACCESSBLOCKSIZE
            static constexpr std::size_t accessblocksize = MALLOCMC_CP_SCATTER_ACCESSBLOCKSIZE;
#else
            static constexpr std::size_t accessblocksize = HeapProperties::accessblocksize;
#endif
#ifdef MALLOCMC_CP_SCATTER_REGIONSIZE
            static constexpr std::uint32_t regionsize = MALLOCMC_CP_SCATTER_REGIONSIZE;
#else
            static constexpr std::uint32_t regionsize = HeapProperties::regionsize;
#endif
#ifdef MALLOCMC_CP_SCATTER_WASTEFACTOR
            static constexpr std::uint32_t wastefactor = MALLOCMC_CP_SCATTER_WASTEFACTOR;
#else
            static constexpr std::uint32_t wastefactor = HeapProperties::wastefactor;
#endif
#ifdef MALLOCMC_CP_SCATTER_RESETFREEDPAGES
            static constexpr bool resetfreedpages = MALLOCMC_CP_SCATTER_RESETFREEDPAGES;
#else
            static constexpr bool resetfreedpages = HeapProperties::resetfreedpages;
#endif
#ifdef MALLOCMC_CP_SCATTER_HASHINGK
            static constexpr std::uint32_t hashingK = MALLOCMC_CP_SCATTER_HASHINGK;
#else
            static constexpr std::uint32_t hashingK = HashingProperties::hashingK;
#endif
#ifdef MALLOCMC_CP_SCATTER_HASHINGDISTMP
            static constexpr std::uint32_t hashingDistMP = MALLOCMC_CP_SCATTER_HASHINGDISTMP;
#else
            static constexpr std::uint32_t hashingDistMP = HashingProperties::hashingDistMP;
#endif
#ifdef MALLOCMC_CP_SCATTER_HASHINGDISTWP
            static constexpr std::uint32_t hashingDistWP = MALLOCMC_CP_SCATTER_HASHINGDISTWP;
#else
            static constexpr std::uint32_t hashingDistWP = HashingProperties::hashingDistWP;
#endif
#ifdef MALLOCMC_CP_SCATTER_HASHINGDISTWPREL
            static constexpr std::uint32_t hashingDistWPRel = MALLOCMC_CP_SCATTER_HASHINGDISTWPREL;
#else
            static constexpr std::uint32_t hashingDistWPRel = HashingProperties::hashingDistWPRel;
#endif

            static constexpr std::uint32_t _pagesize = pagesize;
            static constexpr std::size_t _accessblocksize = accessblocksize;
            static constexpr std::uint32_t _regionsize = regionsize;
            static constexpr std::uint32_t _wastefactor = wastefactor;
            static constexpr bool _resetfreedpages = resetfreedpages;

            //! Derived constants (Scatter.hpp:173-187).
            static constexpr std::uint32_t HierarchyThreshold = (pagesize - 2u * 4u) / 33u;
            static constexpr std::uint32_t minChunkSize = (pagesize - 32u * 4u) / (32u * 32u);

            //! Device view of the heap (the reference keeps _page, _ptes,
            //! _regions, _firstfreeblock ... as members, Scatter.hpp:260-275);
            //! filled by initHeap.
            detail::ScatterHeap heap{};

            //! create (Scatter.hpp:895-912): a chunk for requests up to a
            //! page, a run of pages above it, nullptr for 0 bytes or when
            //! the heap is full.  AlignmentPolicy is the heap's own: its
            //! dataAlignment was fixed at initHeap.
            template<typename AlignmentPolicy, typename Acc>
            MALLOCMC_D void* create(const Acc&, std::uint32_t bytes) const
            {
                return heap.template create<detail::WaveIds>(bytes);
            }

            //! destroy (Scatter.hpp:918-946); nullptr is ignored.
            template<typename Acc>
            MALLOCMC_D void destroy(const Acc&, void* mem) const
            {
                heap.template destroy<detail::WaveIds>(mem);
            }

            //! isOOM (Scatter.hpp:1051-1055): a lane asked for memory and got none.
            MALLOCMC_HD static bool isOOM(void* p, std::size_t s)
            {
                return s != 0u && p == nullptr;
            }

            /** getAvailableSlotsAccelerator (Scatter.hpp:1311-1352): the active
             * lanes of the calling wavefront count the chunks of slotSize
             * (already padded) the heap could still hand out, splitting the
             * pages between them, and each lane receives the total.  Every
             * active lane of the wave must call it; the reference asks the
             * same of the active threads of a warp. */
            template<typename AlignmentPolicy, typename Acc>
            MALLOCMC_D unsigned getAvailableSlotsAccelerator(const Acc&, std::size_t slotSize) const
            {
                const std::uint32_t s = static_cast<std::uint32_t>(slotSize);
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

            /** initHeap (Scatter.hpp:1057-1150): lays the heap out in `pool`
             * (memsize bytes) on `device`, enqueuing its init kernel on
             * `queue`, and fills devAllocator->heap.  The distribution flag
             * stored with the heap (which batch kernels use) follows the
             * allocator's DistributionPolicy; the data alignment, its
             * AlignmentPolicy.  Throws std::invalid_argument for a null pool
             * with memsize != 0 (as the reference does) or a geometry the
             * heap cannot take.  Returns the C heap that owns the metadata
             * (mmc_destroy releases it; the pool stays the caller's). */
            template<typename Acc, typename T_DeviceAllocator>
            static mmc_heap* initHeap(int device, hipStream_t queue, T_DeviceAllocator* devAllocator, void* pool,
                                      std::size_t memsize)
            {
                if(pool == nullptr && memsize != 0u)
                    throw std::invalid_argument("Scatter policy cannot use nullptr for non-empty memory pools. "
                                                "Maybe you are using an incompatible ReservePoolPolicy or "
                                                "AlignmentPolicy.");
                if(hipSetDevice(device) != hipSuccess)
                    throw std::runtime_error("hipSetDevice failed");
                using Dist = typename T_DeviceAllocator::DistributionPolicy;
                using Align = typename T_DeviceAllocator::AlignmentPolicy;
                const detail::ScatterConfig r = runtimeConfig(static_cast<std::uint32_t>(Align::dataAlignment));
                mmc_config c{};
                c.pagesize = r.pagesize;
                c.regionsize = r.regionsize;
                c.accessblocksize = r.accessblocksize;
                c.wastefactor = r.wastefactor;
                c.resetfreedpages = r.resetfreedpages;
                c.hashingK = r.hashingK;
                c.hashingDistMP = r.hashingDistMP;
                c.hashingDistWP = r.hashingDistWP;
                c.hashingDistWPRel = r.hashingDistWPRel;
                c.dataAlignment = r.dataAlignment;
                c.distribution = detail::coalescesOf<Dist>() ? MMC_DIST_XMALLOCSIMD : MMC_DIST_NOOP;
                c.flags = 0u;
                mmc_heap* h = nullptr;
                check(mmc_create_in_pool(&c, pool, memsize, queue, &h), "initHeap");
                std::uint64_t n = sizeof(detail::ScatterHeap);
                const int rc = mmc_get_handle(h, &devAllocator->heap, &n);
                if(rc != MMC_OK)
                {
                    mmc_destroy(h);
                    check(rc, "initHeap");
                }
                return h;
            }

            /** getAvailableSlotsHost (Scatter.hpp:1235-1306): counts the chunks
             * of slotSize (already padded) on `queue` and blocks until the
             * count is back. */
            template<typename Acc>
            static unsigned getAvailableSlotsHost(int device, hipStream_t queue, std::size_t slotSize, mmc_heap* h)
            {
                (void) device;
                std::uint32_t out = 0u;
                check(mmc_available_slots(h, slotSize, queue, &out), "getAvailableSlots");
                return out;
            }

            //! Runtime configuration for the heap; `dataAlignment` comes from the
            //! alignment policy.
            static detail::ScatterConfig runtimeConfig(std::uint32_t dataAlignment)
            {
                detail::ScatterConfig c;
                c.pagesize = pagesize;
                c.regionsize = regionsize;
                c.accessblocksize = accessblocksize;
                c.wastefactor = wastefactor;
                c.resetfreedpages = resetfreedpages ? 1u : 0u;
                c.hashingK = hashingK;
                c.hashingDistMP = hashingDistMP;
                c.hashingDistWP = hashingDistWP;
                c.hashingDistWPRel = hashingDistWPRel;
                c.dataAlignment = dataAlignment;
                return c;
            }

            //! Host errors as exceptions, as the reference's host functions raise them.
            static void check(int rc, const char* what)
            {
                if(rc == MMC_OK)
                    return;
                const std::string msg = std::string(what) + ": " + mmc_last_error();
                if(rc == MMC_EINVAL)
                    throw std::invalid_argument(msg);
                throw std::runtime_error(msg);
            }

            //! Scatter::classname (Scatter.hpp:1354-1370)
            static std::string classname()
            {
                std::stringstream ss;
                ss << "Scatter[";
                ss << "pagesize=" << pagesize << ",";
                ss << "accessblocksize=" << accessblocksize << ",";
                ss << "regionsize=" << regionsize << ",";
                ss << "wastefactor=" << wastefactor << ",";
                ss << "resetfreedpages=" << resetfreedpages << ",";
                ss << "minChunkSize=" << minChunkSize << ",";
                ss << "HierarchyThreshold=" << HierarchyThreshold << ",";
                ss << "hashingK=" << hashingK << ",";
                ss << "hashingDistMP=" << hashingDistMP << ",";
                ss << "hashingDistWP=" << hashingDistWP << ",";
                ss << "hashingDistWPRel=" << hashingDistWPRel << "]";
                return ss.str();
            }
        };
    } // namespace CreationPolicies
} // namespace mallocMC
