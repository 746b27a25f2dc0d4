/*
 * CreationPolicies::Scatter: the ScatterAlloc heap policy with the
 * reference's configuration surface (src/include/mallocMC/creationPolicies/Scatter.hpp:56-207):
 * compile-time config structs, overridable by -D MALLOCMC_CP_SCATTER_* with
 * the precedence default struct < template struct < command-line macro.  The
 * values become the runtime geometry of detail::ScatterHeap, whose device
 * functions implement create/destroy/getAvailableSlots.
 */
#pragma once

#include "../detail/scatter_heap.hpp"

#include <cstddef>
#include <cstdint>
#include <sstream>
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
