/*
 * ScatterAlloc heap core for gfx950: geometry, metadata layout and the
 * per-lane allocation protocol of mallocMC's Scatter creation policy
 * (src/include/mallocMC/creationPolicies/Scatter.hpp), executed with runtime
 * geometry so one build serves every pagesize/regionsize/hash setting.
 *
 * Heap layout inside the pool (identical to Scatter.hpp:999-1014 so that heap
 * offsets are bit-exact): numpages pages, then one 12-byte PTE per page
 * {chunksize, count, bitmask}, then one uint32 fill level per region.  The
 * three mutable heap scalars (_firstFreePageBased, _firstfreeblock,
 * _pagebasedMutex, Scatter.hpp:269-271) live in a separate device block, one
 * 128-byte L2 line each.
 *
 * The identity policy (`Ids`) supplies lane / wave / CU ids.  `WaveIds` reads
 * the hardware; `SerialIds` pins them to the values of the reference's CPU
 * backend (warpSize 1, all ids 0, mallocMC_utils.hpp:85-87), which defines
 * the serialized request stream whose offsets must match the oracle.
 *
 * Divergences from the reference (undefined or unsafe there):
 *   D1 heap < accessblocksize: number of access blocks clamped to 1 (the
 *      reference divides by zero, Scatter.hpp:992-993).
 *   D2 the second page-based scan treats index numpages as "not free" (the
 *      reference reads one PTE past the end, Scatter.hpp:785-788, 835).
 *   D3 free of a pointer whose page has chunksize 0 is ignored (division by
 *      zero at Scatter.hpp:929).
 *   D4 a page emptied by concurrent frees is always reset (see deallocChunked).
 *   D5 a lane that formats a page during its concurrent reset waits for the
 *      reset and then uses the page (see tryUsePage).
 *   D4 and D5 make the same state transitions as the reference on serialized
 *   streams; they close races the reference leaves open under concurrency.
 *   Native fix (flag kFixPagesizeLeak, off by default): a chunk of exactly
 *      pagesize is released through the chunked path, so its page does not
 *      leak one slot (Scatter.hpp:906-908 vs :942-945).
 */
#pragma once

#include "../mallocMC_utils.hpp"

#include <cstdint>

namespace mallocMC
{
    namespace detail
    {
        using u32 = std::uint32_t;
        using u64 = std::uint64_t;

        constexpr u64 kNullOffset = ~0ull;

        enum ScatterFlags : u32
        {
            kFixPagesizeLeak = 1u << 0,
        };

        //! Upper bound on the sleeps a lane spends waiting for a concurrent
        //! page reset to unlock (about 30 ms); past it the lane gives up on
        //! that page instead of waiting forever on a count that never drops.
        constexpr u32 kResetWaitLimit = 1u << 20;

        //! The three mutable heap scalars, each on its own 128-byte L2 line.
        struct alignas(128) ScatterScalars
        {
            u32 firstFreePageBased;
            u32 pad0[31];
            u32 firstfreeblock;
            u32 pad1[31];
            u32 pagebasedMutex;
            u32 pad2[31];
        };

        //! Runtime heap configuration (Scatter.hpp:56-85 + Shrink.hpp:60-66).
        struct ScatterConfig
        {
            u32 pagesize = 4096u;
            u32 regionsize = 16u;
            u64 accessblocksize = 2ull * 1024u * 1024u * 1024u;
            u32 wastefactor = 2u;
            u32 resetfreedpages = 0u;
            u32 hashingK = 38183u;
            u32 hashingDistMP = 17497u;
            u32 hashingDistWP = 1u;
            u32 hashingDistWPRel = 1u;
            u32 dataAlignment = 16u;
            u32 flags = 0u;
        };

        //! Geometry computed by the init kernel (Scatter.hpp:962-1014).
        struct ScatterGeometry
        {
            u32 numpages = 0;
            u32 numregions = 0;
            u32 accessblocks = 0;
            u32 pagesperblock = 0;
            u32 hierarchyThreshold = 0;
            u32 minChunkSize = 0;
            u32 maxOnPageMasks = 0;
            //! 32-bit segment-mask words per page in the side mask array
            //! (at most 32: one 128-byte line per page).
            u32 sideMaskWords = 0;
            u32 clampedBlocks = 0;
            u64 ptesOffset = 0;
            u64 regionsOffset = 0;
            u64 usedBytes = 0;
        };

        MALLOCMC_HD u32 umin32(u32 a, u32 b)
        {
            return a < b ? a : b;
        }
        MALLOCMC_HD u32 umax32(u32 a, u32 b)
        {
            return a > b ? a : b;
        }

        //! Returns nullptr if the configuration is usable, else a message.
        inline const char* validateConfig(const ScatterConfig& c)
        {
            if(c.pagesize < 128u || c.pagesize > (1u << 30))
                return "pagesize must be in [128, 2^30]";
            if(c.regionsize == 0u)
                return "regionsize must be > 0";
            if(c.accessblocksize == 0u)
                return "accessblocksize must be > 0";
            if(c.dataAlignment == 0u || (c.dataAlignment & (c.dataAlignment - 1u)) != 0u)
                return "dataAlignment must be a power of two";
            if(c.wastefactor == 0u)
                return "wastefactor must be > 0";
            return nullptr;
        }

        //! The geometry constants that depend on the pagesize and the data
        //! alignment only: the derived constants of Scatter.hpp:173-187 and
        //! the side-mask width.  computeGeometry uses them, and so do the
        //! compile-time copies of PinnedGeometry.
        struct PageConstants
        {
            u32 hierarchyThreshold;
            u32 minChunkSize;
            u32 maxOnPageMasks;
            u32 sideMaskWords;
        };

        MALLOCMC_HD constexpr PageConstants pageConstants(u32 P, u32 dataAlignment)
        {
            const u32 threshold = (P - 2u * 4u) / 33u;
            const u32 minChunk = (P - 32u * 4u) / (32u * 32u);
            const u32 minSegment = 32u * minChunk + 4u;
            const u32 onPage = minChunk > threshold ? 0u : (P + (minSegment - 1u)) / minSegment;
            // segments of the smallest chunk a page can hold bound the mask
            // words any page needs (fewer for every larger chunk); padded
            // requests are multiples of dataAlignment
            const u32 align = dataAlignment - 1u;
            const u32 padded = (minChunk + align) & ~align;
            const u32 atLeast = dataAlignment > 1u ? dataAlignment : 1u;
            const u32 cmin = padded > atLeast ? padded : atLeast;
            const u32 side = P / (32u * cmin + 4u) + 1u;
            return PageConstants{threshold, minChunk, onPage < 32u ? onPage : 32u, side < 32u ? side : 32u};
        }

        //! Geometry as initDeviceFunction computes it (Scatter.hpp:962-1014) plus
        //! the derived constants of Scatter.hpp:173-187.
        inline ScatterGeometry computeGeometry(const ScatterConfig& c, u64 memsize)
        {
            ScatterGeometry g;
            const u32 P = c.pagesize;
            const PageConstants pc = pageConstants(P, c.dataAlignment);
            g.hierarchyThreshold = pc.hierarchyThreshold;
            g.minChunkSize = pc.minChunkSize;
            g.maxOnPageMasks = pc.maxOnPageMasks;
            g.sideMaskWords = pc.sideMaskWords;

            const u64 R = c.regionsize;
            u32 numregions = static_cast<u32>(memsize / (R * (12u + static_cast<u64>(P)) + 4u));
            u32 numpages = numregions * static_cast<u32>(R);
            const u64 memorysize = static_cast<u64>(numpages) * P;
            u64 numblocks = memorysize / c.accessblocksize;
            if(numblocks == 0u)
            {
                numblocks = 1u;
                g.clampedBlocks = 1u;
            }
            const u64 memoryperblock = memorysize / numblocks;
            const u64 pagesperblock = memoryperblock / P;
            const u64 regionsperblock = pagesperblock / R;
            numregions = static_cast<u32>(numblocks * regionsperblock);
            numpages = numregions * static_cast<u32>(R);
            u64 ptes = static_cast<u64>(numpages) * P;
            u64 regions = ptes + 12u * static_cast<u64>(numpages);
            if(regions + 4u * static_cast<u64>(numregions) > memsize)
            {
                --numregions;
                numpages = umin32(numregions * static_cast<u32>(R), numpages);
                ptes = static_cast<u64>(numpages) * P;
                regions = ptes + 12u * static_cast<u64>(numpages);
            }
            g.numpages = numpages;
            g.numregions = numregions;
            g.accessblocks = static_cast<u32>(numblocks);
            g.pagesperblock = g.accessblocks ? numpages / g.accessblocks : 0u;
            g.ptesOffset = ptes;
            g.regionsOffset = regions;
            g.usedBytes = regions + 4u * static_cast<u64>(numregions);
            return g;
        }

        //! Lane / wave / CU identities of the executing lane.
        struct WaveIds
        {
            static constexpr u32 warpSize = 64u;
            static MALLOCMC_D u32 lane()
            {
                return hw::laneid();
            }
            static MALLOCMC_D u32 warp()
            {
                return hw::waveslot();
            }
            static MALLOCMC_D u32 sm()
            {
                return hw::cuid();
            }
            static MALLOCMC_D u64 active()
            {
                return hw::activemask();
            }
            static MALLOCMC_D u64 lanemaskLt()
            {
                return hw::lanemask_lt();
            }
        };

        //! Identities of the reference's CPU backend: defines the serialized stream.
        struct SerialIds
        {
            static constexpr u32 warpSize = 1u;
            static MALLOCMC_D u32 lane()
            {
                return 0u;
            }
            static MALLOCMC_D u32 warp()
            {
                return 0u;
            }
            static MALLOCMC_D u32 sm()
            {
                return 0u;
            }
            static MALLOCMC_D u64 active()
            {
                return 1u;
            }
            static MALLOCMC_D u64 lanemaskLt()
            {
                return 0u;
            }
        };

        /**
         * By-value view of one heap: what a kernel needs to allocate and free.
         * Trivially copyable; passed to kernels as an argument so the geometry
         * sits in scalar registers.
         */
        struct ScatterHeap
        {
            char* pages = nullptr;
            u32* ptes = nullptr;
            u32* regions = nullptr;
            u32* masks = nullptr; // side segment masks, maskStride words per page
            ScatterScalars* scalars = nullptr;
            u64 memsize = 0;
            u32 pagesize = 0;
            u32 pageShift = 0; // log2(pagesize) if a power of two, else 0
            u32 regionsize = 0;
            u32 wastefactor = 0;
            u32 resetfreedpages = 0;
            u32 hierarchyThreshold = 0;
            u32 minChunkSize = 0;
            u32 maxOnPageMasks = 0;
            u32 maskStride = 0;
            u32 paddedMinChunk = 0;
            u32 alignMask = 0;
            u32 hashingK = 0;
            u32 hashingDistMP = 0;
            u32 hashingDistWP = 0;
            u32 hashingDistWPRel = 0;
            u32 numpages = 0;
            u32 numregions = 0;
            u32 accessblocks = 0;
            u32 pagesperblock = 0;
            u32 flags = 0;
            //! divisions by heap constants: regionsize, numregions, pages per
            //! access block, pagesize
            FastDiv divR;
            FastDiv divRegions;
            FastDiv divPpb;
            FastDiv divP;
            FastDiv divSlice8; //!< pagesperblock / 8: one XCD's slice of an access block

            // ------------------------------------------------ addressing
            // A PTE is 12 bytes (chunk size, count, bitmask) as in the
            // reference.  Chunk size and count share one 8-byte-aligned word
            // so that one 64-bit atomic reads both at the same instant: even
            // pages hold [chunksize, count, bitmask], odd pages (whose entry
            // starts 4 bytes past an 8-byte boundary) [bitmask, chunksize,
            // count].  The PTE array must start 8-byte aligned.
            MALLOCMC_D u32* pteChunksize(u32 p) const
            {
                return ptes + 3u * static_cast<u64>(p) + (p & 1u);
            }
            MALLOCMC_D u32* pteCount(u32 p) const
            {
                return ptes + 3u * static_cast<u64>(p) + 1u + (p & 1u);
            }
            MALLOCMC_D u32* pteBitmask(u32 p) const
            {
                return ptes + 3u * static_cast<u64>(p) + ((p & 1u) != 0u ? 0u : 2u);
            }
            //! Chunk size (low half) and count (high half) of page p.
            MALLOCMC_D u64* ptePair(u32 p) const
            {
                return reinterpret_cast<u64*>(pteChunksize(p));
            }
            MALLOCMC_D char* pageData(u32 p) const
            {
                return pages + static_cast<u64>(p) * pagesize;
            }
            MALLOCMC_D u32 pageOf(u64 offset) const
            {
                return static_cast<u32>(pageShift ? (offset >> pageShift) : (offset / pagesize));
            }
            //! Segment masks of a hierarchical page: word s is segment s's
            //! 32-chunk mask.  The reference keeps them in the page's last
            //! 4*nMasks bytes (onPageMasksPosition, Scatter.hpp:315-318); here
            //! they sit in a side array, one 128-byte line per page, so no
            //! payload byte ever aliases allocator metadata.  Chunk placement
            //! is unchanged: chunks still end where the on-page masks would begin.
            MALLOCMC_D u32* onPageMasks(u32 p, u32 /*nMasks*/) const
            {
                return masks + static_cast<u64>(p) * maskStride;
            }
            //! Shrink::applyPadding (Shrink.hpp:132-137)
            MALLOCMC_HD u32 pad(u32 bytes) const
            {
                return (bytes + alignMask) & ~alignMask;
            }

            // --------------------------------------------- bit utilities
            //! nextspot (Scatter.hpp:293-304)
            static MALLOCMC_D u32 nextspot(u32 bitfield, u32 spot, u32 spots)
            {
                const u32 low = (spot + 1u) == 32u ? 0u : (bitfield >> (spot + 1u));
                const u32 high = bitfield << (spots - (spot + 1u));
                const u32 sel = spots == 32u ? ~0u : ((1u << spots) - 1u);
                const u32 rot = (high | low) & sel;
                const u32 step = hw::ffs(static_cast<u32>(~rot));
                return (spot + step) % spots;
            }

            //! usespot (Scatter.hpp:329-347): claim one free bit, -1 if full.
            template<class Ids>
            static MALLOCMC_D int usespot(u32* bitfield, u32 spots)
            {
                u32 spot = Ids::lane() % spots;
                for(;;)
                {
                    const u32 mask = 1u << spot;
                    const u32 old = atom::or_(bitfield, mask);
                    if((old & mask) == 0u)
                        return static_cast<int>(spot);
                    if(hw::popc(old) >= spots)
                        return -1;
                    spot = nextspot(old, spot, spots);
                }
            }

            //! n / d for n <= pagesize: the float-reciprocal division where it is
            //! exact for the quotients used here (all clamped to 32 or less).
            MALLOCMC_D u32 pdiv(u32 n, u32 d) const
            {
                return pagesize < (1u << 24) ? udiv24(n, d) : n / d;
            }

            //! calcAdditionalChunks (Scatter.hpp:365-383), signed reading:
            //! max(0, pagesize - full * segmentsize - 4) / chunksize, at most 31.
            MALLOCMC_D u32 calcAdditional(u32 fullsegments, u32 segmentsize, u32 chunksize) const
            {
                if(fullsegments == 32u)
                    return 0u;
                const u32 used = fullsegments * segmentsize + 4u; // <= pagesize + 4
                if(used >= pagesize)
                    return 0u;
                return umin32(pdiv(pagesize - used, chunksize), 31u);
            }

            //! Number of chunks a page of the given chunk size holds (tryUsePage :494-510).
            MALLOCMC_D u32 pageCapacity(u32 chunksize) const
            {
                if(chunksize <= hierarchyThreshold)
                {
                    const u32 segmentsize = chunksize * 32u + 4u;
                    const u32 full = umin32(32u, pdiv(pagesize, segmentsize));
                    return full * 32u + calcAdditional(full, segmentsize, chunksize);
                }
                return umin32(pdiv(pagesize, chunksize), 32u);
            }

            // ------------------------------------------- chunked path
            //! addChunkHierarchy (Scatter.hpp:397-421)
            template<class Ids>
            MALLOCMC_D char* addChunkHierarchy(u32 chunksize, u32 fullsegments, u32 additional, u32 page) const
            {
                const u32 segments = fullsegments + (additional > 0u ? 1u : 0u);
                u32 spot = Ids::lane() % segments;
                const u32 mask = atom::load(pteBitmask(page));
                if((mask & (1u << spot)) != 0u)
                    spot = nextspot(mask, spot, segments);
                const u32 tries = segments - hw::popc(mask);
                u32* masks = onPageMasks(page, segments);
                for(u32 i = 0; i < tries; ++i)
                {
                    const int hspot = usespot<Ids>(&masks[spot], spot < fullsegments ? 32u : additional);
                    if(hspot != -1)
                        return pageData(page) + static_cast<u64>(32u * spot + static_cast<u32>(hspot)) * chunksize;
                    atom::or_(pteBitmask(page), 1u << spot);
                    spot = nextspot(mask, spot, segments);
                }
                return nullptr;
            }

            //! addChunkNoHierarchy (Scatter.hpp:432-443)
            template<class Ids>
            MALLOCMC_D char* addChunkNoHierarchy(u32 chunksize, u32 page, u32 spots) const
            {
                const int spot = usespot<Ids>(pteBitmask(page), spots);
                if(spot == -1)
                    return nullptr;
                return pageData(page) + static_cast<u64>(static_cast<u32>(spot)) * chunksize;
            }

            //! tryUsePage (Scatter.hpp:454-518); range check of allocChunked (:589-591).
            //! `formatted`: this lane set the page's chunk size from 0 just before
            //! (divergence D5 below).
            template<class Ids>
            MALLOCMC_D char* tryUsePage(u32 page, u32 chunksize, u32 bytes, u32 maxchunksize, bool formatted) const
            {
                u32 waited = 0u;
                for(;;)
                {
                    char* chunk = nullptr;
                    const u32 filllevel = atom::add(pteCount(page), 1u);
                    bool tryAlloc = !resetfreedpages;
                    if(resetfreedpages && filllevel < pagesize)
                    {
                        const u32 oldcs = atom::cas(pteChunksize(page), 0u, chunksize);
                        if(oldcs == 0u || (oldcs >= bytes && oldcs <= maxchunksize))
                            tryAlloc = true;
                        if(oldcs != 0u)
                            chunksize = oldcs;
                    }
                    if(tryAlloc)
                    {
                        if(chunksize <= hierarchyThreshold)
                        {
                            const u32 segmentsize = chunksize * 32u + 4u;
                            const u32 full = umin32(32u, pagesize / segmentsize);
                            const u32 additional = calcAdditional(full, segmentsize, chunksize);
                            if(filllevel < full * 32u + additional)
                                chunk = addChunkHierarchy<Ids>(chunksize, full, additional, page);
                        }
                        else
                        {
                            const u32 chunksinpage = umin32(pagesize / chunksize, 32u);
                            if(filllevel < chunksinpage)
                                chunk = addChunkNoHierarchy<Ids>(chunksize, page, chunksinpage);
                        }
                    }
                    if(chunk == nullptr)
                        atom::sub(pteCount(page), 1u);
                    // D5: the lane formatted a page whose reset (count locked at
                    // pagesize, Scatter.hpp:686-711) was still running.  Leaving
                    // would strand the page formatted but empty; the reset is a
                    // short straight-line section, so wait for it and use the page.
                    // Never taken on a serialized stream (count stays < pagesize).
                    if(chunk == nullptr && formatted && resetfreedpages && filllevel >= pagesize
                       && waited < kResetWaitLimit)
                    {
                        while(atom::load(pteCount(page)) >= pagesize && waited < kResetWaitLimit)
                        {
                            __builtin_amdgcn_s_sleep(1);
                            ++waited;
                        }
                        ++waited;
                        continue;
                    }
                    return chunk;
                }
            }

            //! allocChunked (Scatter.hpp:527-637)
            template<class Ids>
            MALLOCMC_D char* allocChunked(u32 bytes) const
            {
                const u32 P = pagesize;
                const u32 R = regionsize;
                const u32 minAllocation = umax32(bytes, paddedMinChunk);
                const u32 reloff = Ids::warpSize * minAllocation / P;
                const u32 startInBlock
                    = (minAllocation * hashingK + hashingDistMP * Ids::sm()
                       + (hashingDistWP + hashingDistWPRel * reloff) * Ids::warp())
                    % pagesperblock;
                const u32 maxchunksize = umin32(P, umax32(wastefactor * bytes, paddedMinChunk));
                const u32 globalStart = startInBlock + atom::load(&scalars->firstfreeblock) * pagesperblock;

                u32 checklevel = R * 3u / 4u;
                for(u32 finder = 0; finder < 2u; ++finder)
                {
                    u32 gp = globalStart;
                    do
                    {
                        const u32 region = gp / R;
                        const u32 fill = atom::load(&regions[region]);
                        const u32 regionOffset = region * R;
                        if(fill < checklevel)
                        {
                            u32 pir = gp;
                            do
                            {
                                const u32 before = atom::cas(pteChunksize(pir), 0u, minAllocation);
                                u32 use = 0u;
                                if(before == 0u)
                                    use = minAllocation;
                                else if(before >= bytes && before <= maxchunksize)
                                    use = before;
                                if(use != 0u)
                                {
                                    char* res = tryUsePage<Ids>(pir, use, bytes, maxchunksize, before == 0u);
                                    if(res != nullptr)
                                        return res;
                                }
                                pir = regionOffset + ((pir + 1u) % R);
                            } while(pir != gp);
                            if(fill + 1u <= R)
                                atom::cas(&regions[region], fill, fill + 1u);
                        }
                        gp = (gp + R) % numpages;
                        if(gp % pagesperblock == 0u)
                        {
                            const u32 abid = gp / pagesperblock;
                            if(abid > atom::load(&scalars->firstfreeblock))
                                atom::store(&scalars->firstfreeblock, abid);
                        }
                    } while(gp != globalStart);
                    checklevel = R + 1u;
                }
                return nullptr;
            }

            //! Page::init (Scatter.hpp:249-258): clear the on-page mask area.
            //! Clears a page's segment masks (Page::init, Scatter.hpp:249-258;
            //! the reference calls it when a page-based allocation is freed,
            //! :879-881).  Returning exchanges: the caller may consume the
            //! result to order later atomics after these have been performed.
            MALLOCMC_D u32 pageInit(u32 page) const
            {
                u32* w = masks + static_cast<u64>(page) * maskStride;
                u32 acc = 0u;
                for(u32 i = 0; i < maskStride; ++i)
                    acc |= atom::exch(&w[i], 0u);
                return acc;
            }

            //! deallocChunked (Scatter.hpp:646-736)
            template<class Ids>
            MALLOCMC_D void deallocChunked(char* mem, u32 page, u32 chunksize) const
            {
                const u32 P = pagesize;
                const u32 inpage = static_cast<u32>(mem - pageData(page));
                if(chunksize <= hierarchyThreshold)
                {
                    const u32 segmentsize = chunksize * 32u + 4u;
                    const u32 full = umin32(32u, P / segmentsize);
                    const u32 additional = calcAdditional(full, segmentsize, chunksize);
                    const u32 segment = inpage / (chunksize * 32u);
                    const u32 within = (inpage - segment * (chunksize * 32u)) / chunksize;
                    const u32 nMasks = full + (additional > 0u ? 1u : 0u);
                    u32* masks = onPageMasks(page, nMasks);
                    atom::and_(&masks[segment], ~(1u << within));
                    atom::and_(pteBitmask(page), ~(1u << segment));
                }
                else
                {
                    const u32 segment = inpage / chunksize;
                    atom::and_(pteBitmask(page), ~(1u << segment));
                }

                // D4: the reference locks with CAS(count, 1 -> P) and decrements
                // otherwise (Scatter.hpp:686-716); two concurrent frees can both
                // miss the lock and leave the empty page formatted.  Here the lane
                // whose decrement empties the page takes the lock with
                // CAS(count, 0 -> P).  On a serialized stream both forms make the
                // same state transitions and return the same old fill level.
                // The reference also clears the page's on-page masks here
                // (Page::init, Scatter.hpp:249-258, 691): they share the page
                // with payload, which a flat chunk size may have overwritten.
                // The side masks hold no payload, and every claimed bit is
                // cleared by its free before the count drops, so they are
                // already clear when the count reaches zero; clearing them
                // again could wipe a bit a concurrent wave-path claim has just
                // set ahead of its count reservation.
                const u32 oldfill = atom::sub(pteCount(page), 1u);
                if(resetfreedpages && oldfill == 1u && atom::cas(pteCount(page), 0u, P) == 0u)
                {
                    atom::cas(pteChunksize(page), chunksize, 0u);
                    atom::fence_device();
                    atom::sub(pteCount(page), P);
                }

                if(oldfill == P / 2u / chunksize)
                {
                    const u32 region = page / regionsize;
                    atom::exch(&regions[region], 0u);
                    const u32 block = page / pagesperblock;
                    if(Ids::warp() + Ids::lane() == 0u)
                        atom::min(&scalars->firstfreeblock, block);
                }
            }

            // ---------------------------------------- page-based path
            //! markpages (Scatter.hpp:745-764)
            MALLOCMC_D bool markpages(u32 startpage, u32 npages, u32 bytes) const
            {
                u32 abortAt = 0xffffffffu;
                for(u32 p = startpage; p < startpage + npages; ++p)
                {
                    if(atom::cas(pteChunksize(p), 0u, bytes) != 0u)
                    {
                        abortAt = p;
                        break;
                    }
                }
                if(abortAt == 0xffffffffu)
                    return true;
                for(u32 p = startpage; p < abortAt; ++p)
                    atom::cas(pteChunksize(p), bytes, 0u);
                return false;
            }

            //! allocPageBasedSingleRegion (Scatter.hpp:775-813), divergence D2.
            MALLOCMC_D char* allocPageBasedSingleRegion(u32 startpage, u32 endpage, u32 bytes) const
            {
                const u32 pagestoalloc = divup(bytes, pagesize);
                u32 freecount = 0;
                bool leftFree = false;
                for(u32 sp = startpage + 1u; sp > endpage;)
                {
                    --sp;
                    const bool isFree = sp < numpages && atom::load(pteChunksize(sp)) == 0u;
                    if(isFree)
                    {
                        if(++freecount == pagestoalloc)
                        {
                            if(markpages(sp, pagestoalloc, bytes))
                            {
                                if(!leftFree)
                                    atom::cas(&scalars->firstFreePageBased, startpage, sp - 1u);
                                return pageData(sp);
                            }
                        }
                    }
                    else
                    {
                        leftFree = true;
                        freecount = 0;
                    }
                }
                return nullptr;
            }

            //! allocPageBasedSingle (Scatter.hpp:824-840): global spin mutex.
            MALLOCMC_D char* allocPageBasedSingle(u32 bytes) const
            {
                while(atom::exch(&scalars->pagebasedMutex, 1u) != 0u)
                    __builtin_amdgcn_s_sleep(2);
                const u32 spage = atom::load(&scalars->firstFreePageBased);
                char* res = allocPageBasedSingleRegion(spage, 0u, bytes);
                if(res == nullptr)
                    res = allocPageBasedSingleRegion(numpages, spage, bytes);
                atom::fence_device();
                atom::exch(&scalars->pagebasedMutex, 0u);
                return res;
            }

            //! allocPageBased (Scatter.hpp:848-866): one active lane at a time.
            template<class Ids>
            MALLOCMC_D char* allocPageBased(u32 bytes) const
            {
                char* res = nullptr;
                const u64 mask = Ids::active();
                const u32 num = hw::popc(mask);
                const u32 localId = hw::popc(Ids::lanemaskLt() & mask);
                for(u32 active = 0; active < num; ++active)
                    if(active == localId)
                        res = allocPageBasedSingle(bytes);
                return res;
            }

            //! deallocPageBased (Scatter.hpp:874-886)
            MALLOCMC_D void deallocPageBased(u32 page, u32 bytes) const
            {
                const u32 npages = divup(bytes, pagesize);
                for(u32 p = page; p < page + npages; ++p)
                    pageInit(p);
                atom::fence_device();
                for(u32 p = page; p < page + npages; ++p)
                    atom::cas(pteChunksize(p), bytes, 0u);
                atom::max(&scalars->firstFreePageBased, page + npages - 1u);
            }

            // ------------------------------------------------- entry
            //! create (Scatter.hpp:895-912); bytes already padded.
            template<class Ids>
            MALLOCMC_D char* create(u32 bytes) const
            {
                if(bytes == 0u)
                    return nullptr;
                if(bytes <= pagesize)
                    return allocChunked<Ids>(bytes);
                return allocPageBased<Ids>(bytes);
            }

            //! destroy (Scatter.hpp:918-946) with divergence D3.
            template<class Ids>
            MALLOCMC_D void destroy(void* memv) const
            {
                if(memv == nullptr)
                    return;
                char* mem = static_cast<char*>(memv);
                const u32 page = pageOf(static_cast<u64>(mem - pages));
                const u32 chunksize = atom::load(pteChunksize(page));
                if(chunksize == 0u)
                    return;
                const u32 inpage = static_cast<u32>(mem - pageData(page));
                const u32 block = inpage / chunksize;
                const u32 inblock = inpage - block * chunksize;
                if(inblock != 0u)
                {
                    u32* counter = reinterpret_cast<u32*>(pageData(page) + static_cast<u64>(block) * chunksize);
                    const u32 old = atom::sub(counter, 1u);
                    if(old != 1u)
                        return;
                    mem = reinterpret_cast<char*>(counter);
                }
                if(chunksize < pagesize || (chunksize == pagesize && (flags & kFixPagesizeLeak)))
                    deallocChunked<Ids>(mem, page, chunksize);
                else
                    deallocPageBased(page, chunksize);
            }

            // ------------------------------------------ slot counting
            //! getAvailaibleSlotsDeviceFunction (Scatter.hpp:1158-1219)
            MALLOCMC_D u32 availableSlotsPartial(u64 slotSize, u32 gid, u32 stride) const
            {
                u32 slots = 0u;
                if(slotSize < pagesize)
                {
                    const u32 maxchunksize = umin32(pagesize, wastefactor * static_cast<u32>(slotSize));
                    for(u32 p = gid; p < numpages; p += stride)
                    {
                        // chunk size and count in one load of their paired
                        // word; free chunks = capacity - count
                        // (countFreeChunksInPage, Scatter.hpp:1112-1139,
                        // reads the count after the chunk size)
                        const u64 pr = atom::load(ptePair(p));
                        const u32 cs = static_cast<u32>(pr);
                        const u32 filled = static_cast<u32>(pr >> 32);
                        if(cs >= slotSize && cs <= maxchunksize)
                            slots += pageCapacity(cs) - filled;
                        else if(cs == 0u)
                            slots += pageCapacity(umax32(static_cast<u32>(slotSize), paddedMinChunk)) - filled;
                    }
                }
                else
                {
                    if(gid > 0u)
                        return 0u;
                    const u32 pagestoalloc = divup(static_cast<u32>(slotSize), pagesize);
                    u32 freecount = 0u;
                    for(u32 p = numpages; p > 0u;)
                    {
                        --p;
                        if(atom::load(pteChunksize(p)) == 0u)
                        {
                            if(++freecount == pagestoalloc)
                            {
                                freecount = 0u;
                                ++slots;
                            }
                        }
                        else
                            freecount = 0u;
                    }
                }
                return slots;
            }

            MALLOCMC_D u64 offsetOf(const void* p) const
            {
                return p == nullptr ? kNullOffset : static_cast<u64>(static_cast<const char*>(p) - pages);
            }
        };

        //! Builds the by-value view for a pool already laid out by computeGeometry.
        inline ScatterHeap makeHeapView(
            const ScatterConfig& c,
            const ScatterGeometry& g,
            void* pool,
            u64 memsize,
            ScatterScalars* scalars,
            u32* sideMasks)
        {
            ScatterHeap h;
            h.pages = static_cast<char*>(pool);
            h.masks = sideMasks;
            h.maskStride = g.sideMaskWords;
            h.ptes = reinterpret_cast<u32*>(h.pages + g.ptesOffset);
            h.regions = reinterpret_cast<u32*>(h.pages + g.regionsOffset);
            h.scalars = scalars;
            h.memsize = memsize;
            h.pagesize = c.pagesize;
            h.pageShift = 0u;
            if((c.pagesize & (c.pagesize - 1u)) == 0u)
            {
                u32 s = 0;
                while((1u << s) < c.pagesize)
                    ++s;
                h.pageShift = s;
            }
            h.regionsize = c.regionsize;
            h.wastefactor = c.wastefactor;
            h.resetfreedpages = c.resetfreedpages ? 1u : 0u;
            h.hierarchyThreshold = g.hierarchyThreshold;
            h.minChunkSize = g.minChunkSize;
            h.maxOnPageMasks = g.maxOnPageMasks;
            h.alignMask = c.dataAlignment - 1u;
            h.paddedMinChunk = (g.minChunkSize + h.alignMask) & ~h.alignMask;
            h.hashingK = c.hashingK;
            h.hashingDistMP = c.hashingDistMP;
            h.hashingDistWP = c.hashingDistWP;
            h.hashingDistWPRel = c.hashingDistWPRel;
            h.numpages = g.numpages;
            h.numregions = g.numregions;
            h.accessblocks = g.accessblocks;
            h.pagesperblock = g.pagesperblock;
            h.flags = c.flags;
            h.divR = FastDiv(c.regionsize);
            h.divRegions = FastDiv(g.numregions);
            h.divPpb = FastDiv(g.pagesperblock);
            h.divP = FastDiv(c.pagesize);
            h.divSlice8 = FastDiv(g.pagesperblock / 8u);
            return h;
        }

        //! log2 of a power of two.
        MALLOCMC_HD constexpr u32 log2Pow2(u32 x)
        {
            u32 s = 0u;
            while((1u << s) < x)
                ++s;
            return s;
        }

        //! A heap view's configuration pinned at compile time: pagesize P and
        //! regionsize R (powers of two), wastefactor 2, data alignment 16 and
        //! the default hashing constants.  A kernel built for it overwrites
        //! those fields of its copy of the view with constants, so that the
        //! compiler folds them (divisions by R and P become shifts, thresholds
        //! and hash factors become immediates) and keeps fewer scalar
        //! registers live.  `matches` tells the launch whether a heap's view
        //! holds exactly these values; the heap-size fields (page and region
        //! counts, access blocks, pointers) and resetfreedpages and flags stay
        //! as the view has them.
        template<u32 P, u32 R>
        struct PinnedGeometry
        {
            static_assert(P >= 128u && (P & (P - 1u)) == 0u && R != 0u && (R & (R - 1u)) == 0u,
                          "pagesize and regionsize are powers of two");
            static constexpr u32 kAlign = 16u;
            static constexpr PageConstants kPage = pageConstants(P, kAlign);

            MALLOCMC_HD static ScatterHeap apply(ScatterHeap h)
            {
                constexpr ScatterConfig kDefaults{};
                h.pagesize = P;
                h.pageShift = log2Pow2(P);
                h.regionsize = R;
                h.wastefactor = 2u;
                h.hierarchyThreshold = kPage.hierarchyThreshold;
                h.minChunkSize = kPage.minChunkSize;
                h.maxOnPageMasks = kPage.maxOnPageMasks;
                h.maskStride = kPage.sideMaskWords;
                h.alignMask = kAlign - 1u;
                h.paddedMinChunk = (kPage.minChunkSize + kAlign - 1u) & ~(kAlign - 1u);
                h.hashingK = kDefaults.hashingK;
                h.hashingDistMP = kDefaults.hashingDistMP;
                h.hashingDistWP = kDefaults.hashingDistWP;
                h.hashingDistWPRel = kDefaults.hashingDistWPRel;
                h.divR.d = R;
                h.divR.m = 0u;
                h.divR.s = log2Pow2(R);
                h.divP.d = P;
                h.divP.m = 0u;
                h.divP.s = log2Pow2(P);
                return h;
            }

            static bool matches(const ScatterHeap& h)
            {
                const ScatterHeap p = apply(h);
                return h.pagesize == p.pagesize && h.pageShift == p.pageShift && h.regionsize == p.regionsize
                    && h.wastefactor == p.wastefactor && h.hierarchyThreshold == p.hierarchyThreshold
                    && h.minChunkSize == p.minChunkSize && h.maxOnPageMasks == p.maxOnPageMasks
                    && h.maskStride == p.maskStride && h.alignMask == p.alignMask
                    && h.paddedMinChunk == p.paddedMinChunk && h.hashingK == p.hashingK
                    && h.hashingDistMP == p.hashingDistMP && h.hashingDistWP == p.hashingDistWP
                    && h.hashingDistWPRel == p.hashingDistWPRel && h.divR.d == p.divR.d && h.divR.m == p.divR.m
                    && h.divR.s == p.divR.s && h.divP.d == p.divP.d && h.divP.m == p.divP.m && h.divP.s == p.divP.s;
            }
        };

        //! No field pinned: the view as it is, for every configuration.
        struct RuntimeGeometry
        {
            MALLOCMC_HD static ScatterHeap apply(const ScatterHeap& h)
            {
                return h;
            }

            static bool matches(const ScatterHeap&)
            {
                return true;
            }
        };

        //! The configurations with kernels of their own (PinnedGeometry): the
        //! pagesizes and regionsizes that lead the one-heap tuning space, with
        //! the default wastefactor, alignment and hashing.
        template<class... G>
        struct GeometryList
        {
        };
        using PinnedGeometries = GeometryList<PinnedGeometry<65536, 32>, PinnedGeometry<65536, 16>,
                                              PinnedGeometry<16384, 16>, PinnedGeometry<16384, 32>>;

        //! The length of the list.
        template<class... G>
        constexpr u32 geometryCount(GeometryList<G...>)
        {
            return static_cast<u32>(sizeof...(G));
        }

        //! How many geometries of the list accept the view.
        template<class... G>
        u32 countMatches(const ScatterHeap& v, GeometryList<G...>)
        {
            return (0u + ... + (G::matches(v) ? 1u : 0u));
        }

        //! Calls launch(G{}) for the first geometry of the list that the view
        //! matches, or launch(RuntimeGeometry{}).
        template<class F>
        void withGeometry(const ScatterHeap&, F&& launch, GeometryList<>)
        {
            launch(RuntimeGeometry{});
        }

        template<class F, class G0, class... Gs>
        void withGeometry(const ScatterHeap& v, F&& launch, GeometryList<G0, Gs...>)
        {
            if(G0::matches(v))
                launch(G0{});
            else
                withGeometry(v, launch, GeometryList<Gs...>{});
        }

        //! Position of the first geometry of the list that every one of the
        //! n views matches; the list's length if none does.
        template<class... G>
        u32 commonGeometry(const ScatterHeap* views, u32 n, GeometryList<G...>)
        {
            auto all = [views, n](auto g) {
                for(u32 i = 0; i < n; ++i)
                    if(!decltype(g)::matches(views[i]))
                        return false;
                return true;
            };
            u32 k = 0u;
            const bool found = ((all(G{}) ? true : (++k, false)) || ...);
            return found ? k : static_cast<u32>(sizeof...(G));
        }

        //! Calls launch(G{}) for geometry `index` of the list, or
        //! launch(RuntimeGeometry{}) for an index past its end.
        template<class F>
        void withGeometryAt(u32, F&& launch, GeometryList<>)
        {
            launch(RuntimeGeometry{});
        }

        template<class F, class G0, class... Gs>
        void withGeometryAt(u32 index, F&& launch, GeometryList<G0, Gs...>)
        {
            if(index == 0u)
                launch(G0{});
            else
                withGeometryAt(index - 1u, launch, GeometryList<Gs...>{});
        }
    } // namespace detail
} // namespace mallocMC
