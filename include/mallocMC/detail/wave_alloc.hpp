/*
 * Wave-cooperative allocation on the Scatter heap (gfx950, wave64).
 *
 * The reference serves every lane on its own (Scatter.hpp:527-637): each lane
 * hashes to a page, CASes the chunk size of every page it visits, reserves
 * one slot with atomicAdd(count) and claims one bit with atomicOr.  Here the
 * 64 lanes of a wavefront cooperate on the same heap metadata:
 *
 *   * lanes are grouped into size classes whose members can share one chunk
 *     size c while respecting the wastefactor rule b <= c <= max(WF*b,
 *     minChunk) of Scatter.hpp:539-545, 589-591 (octaves of the request when
 *     WF >= 2, exact sizes otherwise);
 *   * a region is examined with one coalesced load of its PTEs (lane j reads
 *     page j of the region) instead of one CAS per visited page;
 *   * a page serves the whole group with one atomicAdd(count, k) reservation
 *     and one vector atomicOr per touched 32-chunk mask word; chunk indices
 *     are handed to lanes by ballot rank;
 *   * pages the scan saw unformatted, and flat pages (chunks in the PTE
 *     bitmask), are claimed several at a time with the format, reservation
 *     and bit claim in one round trip; when other waves took the scanned
 *     pages meanwhile, the window is read again;
 *   * frees are grouped by page: one atomicAnd per touched mask word, one
 *     atomicAnd on the segment hints and one atomicSub(count, k).
 *
 * The metadata protocol is the reference's (same PTE fields, segment masks --
 * kept in a side array here --, region fill levels, access-block cursor;
 * a count at or above pagesize is a reset lock and is waited out), so per-lane
 * calls (scatter_heap.hpp) and wave calls can be mixed on one heap.  Two
 * steps use the PTE's paired (chunk size, count) word: a reservation is one
 * 64-bit add that also returns the chunk size, and a wave resets an empty page
 * with one 64-bit CAS instead of the lock / drop / unlock sequence.
 * Requests above a page and coalesced XMallocSIMD slices use the per-lane
 * path.
 *
 * Precondition: every function here is called by all 64 lanes of the wave
 * in uniform control flow (lanes without work pass 0 / nullptr).
 */
#pragma once

#include "device_allocator.hpp"
#include "distribution.hpp"
#include "scatter_heap.hpp"

namespace mallocMC
{
    namespace detail
    {
        //! Reads of one scan window a group may make when its claims find the
        //! window's snapshot stale (pages taken by other waves meanwhile).
        constexpr u32 kWindowReads = 3u;

        //! Tuning knobs of the wave path (autotuned by the batched runner).
        struct WaveTuning
        {
            //! Maximum lanes one group claim serves (1 = per-lane like the reference).
            u32 aggregation = 64u;
            //! Spread start pages by XCD: each XCD hashes into its own slice of
            //! every access block, so its atomics concentrate on its own lines.
            u32 xcdSlices = 0u;
            //! Size classes per octave = 2^classSplit (0..3): finer classes waste
            //! less (chunk <= (1 + 2^-classSplit) x request) but aggregate fewer lanes.
            u32 classSplit = 0u;
            //! Claim on several pages per round trip: flat pages from their
            //! PTE bitmask (claimFlat), unformatted hierarchical pages from
            //! their first segment words (claimFresh, unless freshFast is 0);
            //! 0 claims every page on its own (claimOnPage).
            u32 flatFast = 1u;
            //! After a window without any candidate page, skip ahead 256 pages
            //! per round trip (fastForward).
            u32 fastForward = 1u;
            //! With flatFast: claim unformatted hierarchical pages several per
            //! round trip (claimFresh); 0 formats them one at a time through
            //! claimOnPage.  Many small groups of one size class (the
            //! multi-heap step's 64-slot tiles) can run faster without.
            u32 freshFast = 1u;
        };

        //! The default wave tuning pinned at compile time: a kernel built
        //! for it folds the tuning's branches and constants (the group cap,
        //! the claim forms, no XCD slices) instead of keeping them in
        //! registers; `matches` tells the launch whether a tuning is the
        //! default.
        struct DefaultTuning
        {
            MALLOCMC_HD static WaveTuning apply(const WaveTuning&)
            {
                return WaveTuning{};
            }

            static bool matches(const WaveTuning& t)
            {
                const WaveTuning d{};
                return t.aggregation == d.aggregation && t.xcdSlices == d.xcdSlices && t.classSplit == d.classSplit
                    && t.flatFast == d.flatFast && t.fastForward == d.fastForward && t.freshFast == d.freshFast;
            }
        };

        //! Any tuning, read at run time.
        struct RuntimeTuning
        {
            MALLOCMC_HD static WaveTuning apply(const WaveTuning& t)
            {
                return t;
            }

            static bool matches(const WaveTuning&)
            {
                return true;
            }
        };

        //! Wave-wide reductions and scans.  All 64 lanes must be active.
        //! Reductions use DPP (data-parallel lane moves inside the VALU: quad
        //! butterflies, row rotates and the GFX9 row broadcasts), not LDS
        //! permutes, so a reduction costs a few VALU cycles instead of six
        //! LDS round trips.
        namespace wave
        {
            //! DPP controls (GFX9): quad_perm [1,0,3,2], quad_perm [2,3,0,1],
            //! row_ror:4, row_ror:8, row_bcast:15, row_bcast:31.
            constexpr int kQuadXor1 = 0xB1;
            constexpr int kQuadXor2 = 0x4E;
            constexpr int kRowRor4 = 0x124;
            constexpr int kRowRor8 = 0x128;
            constexpr int kRowBcast15 = 0x142;
            constexpr int kRowBcast31 = 0x143;

            //! DPP move whose lanes outside RowMask, and lanes whose source
            //! lies outside their row, read 0.
            template<int Ctrl, int RowMask = 0xf>
            MALLOCMC_D u32 dppZero(u32 v)
            {
                return static_cast<u32>(
                    __builtin_amdgcn_update_dpp(0, static_cast<int>(v), Ctrl, RowMask, 0xf, true));
            }

            //! Combines all 64 lanes with an associative, commutative op whose
            //! identity is 0 (unsigned max, or, sum); lane 63 ends with the
            //! result, which is returned to all lanes.  Lanes a DPP step does
            //! not feed combine with 0 and keep their value, and lane 63 only
            //! ever reads lanes that hold exact partial results.  With 0 as
            //! the moves' fill, the compiler folds each move into its
            //! combining instruction.
            template<class Op>
            MALLOCMC_D u32 reduce(u32 v, Op op)
            {
                v = op(v, dppZero<kQuadXor1>(v));
                v = op(v, dppZero<kQuadXor2>(v)); // quad value in every lane
                v = op(v, dppZero<kRowRor4>(v));
                v = op(v, dppZero<kRowRor8>(v)); // row (16 lanes) value in every lane
                v = op(v, dppZero<kRowBcast15, 0xa>(v)); // rows 1, 3 add rows 0, 2
                v = op(v, dppZero<kRowBcast31, 0xc>(v)); // rows 2, 3 add lane 31
                return static_cast<u32>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
            }

            MALLOCMC_D u32 max(u32 v)
            {
                return reduce(v, [](u32 a, u32 b) { return a > b ? a : b; });
            }
            MALLOCMC_D u32 min(u32 v)
            {
                return ~reduce(~v, [](u32 a, u32 b) { return a > b ? a : b; });
            }
            MALLOCMC_D u32 bor(u32 v)
            {
                return reduce(v, [](u32 a, u32 b) { return a | b; });
            }
            //! Sum over the 64 lanes (modulo 2^32).
            MALLOCMC_D u32 sum(u32 v)
            {
                return reduce(v, [](u32 a, u32 b) { return a + b; });
            }
            //! DPP controls row_shr:1, 2, 4, 8 (shift within a row of 16 lanes).
            constexpr int kRowShr1 = 0x111;
            constexpr int kRowShr2 = 0x112;
            constexpr int kRowShr4 = 0x114;
            constexpr int kRowShr8 = 0x118;
            //! Exclusive prefix sum over the 64 lanes (modulo 2^32): a
            //! Hillis-Steele scan inside each row of 16 lanes (shifts 1, 2, 4,
            //! 8), then rows 1 and 3 add the last lane of rows 0 and 2, and
            //! rows 2 and 3 add lane 31.  Six DPP additions; the bit-by-bit
            //! form (a ballot and an mbcnt per bit of the value) took about
            //! thirty instructions for values below 64.
            MALLOCMC_D u32 exclusiveSum(u32 v)
            {
                u32 s = v;
                s += dppZero<kRowShr1>(s);
                s += dppZero<kRowShr2>(s);
                s += dppZero<kRowShr4>(s);
                s += dppZero<kRowShr8>(s);
                s += dppZero<kRowBcast15, 0xa>(s);
                s += dppZero<kRowBcast31, 0xc>(s);
                return s - v;
            }
            //! Inclusive prefix sum over the 64 lanes (modulo 2^32).
            MALLOCMC_D u32 inclusiveSum(u32 v)
            {
                return exclusiveSum(v) + v;
            }
            //! Value of lane `l` (l uniform).
            MALLOCMC_D u32 lane(u32 v, u32 l)
            {
                return __builtin_amdgcn_readlane(v, l);
            }
            //! A value every lane holds alike, marked as such: the compiler
            //! cannot always prove it (the claim counts), and a loop that exits
            //! on a value it takes for divergent keeps its live values in
            //! vector registers (16 more in the wave malloc).
            MALLOCMC_D u32 uniform(u32 v)
            {
                return static_cast<u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
            }
            //! Position (0-based) of the r-th set bit of x; x must have > r bits.
            MALLOCMC_D u32 selectBit(u32 x, u32 r)
            {
                u32 pos = 0u;
                for(u32 step = 16u; step > 0u; step >>= 1)
                {
                    const u64 below = (static_cast<u64>(1u) << (pos + step)) - 1u;
                    if(static_cast<u32>(hw::popc(x & below)) <= r)
                        pos += step;
                }
                return pos;
            }
            //! The lowest `n` set bits of x (all of x if it has n or fewer).
            MALLOCMC_D u32 lowestBits(u32 x, u32 n)
            {
                if(n == 0u)
                    return 0u;
                if(static_cast<u32>(hw::popc(x)) <= n)
                    return x;
                if((x & (x + 1u)) == 0u)
                    return (1u << n) - 1u; // x = 0..01..1 (a fresh page's free bits): n < popc(x) <= 32

                const u32 pos = selectBit(x, n - 1u); // position of the n-th set bit
                return pos >= 31u ? x : (x & ((2u << pos) - 1u));
            }
        } // namespace wave

        //! Size class used to group lanes; equal class => one chunk size c
        //! (the class maximum) satisfies b <= c <= max(WF*b, minChunk) for
        //! every member b.  Classes are octaves of the request above the minimum
        //! chunk, each cut into 2^split equal parts.
        MALLOCMC_D u32 sizeClass(const ScatterHeap& h, u32 b, u32 split)
        {
            if(h.paddedMinChunk != 0u && b <= h.paddedMinChunk)
                return 0u;
            if(h.wastefactor < 2u)
                return b;
            const u32 base = h.paddedMinChunk != 0u ? h.paddedMinChunk : 16u;
            const u32 q = h.pdiv(b - 1u, base);
            if(q == 0u)
                return 0x80000000u | b; // below the octave base: exact size
            // octave j: q in [2^j, 2^(j+1)), i.e. b in (base*2^j, base*2^(j+1)]
            const u32 j = 31u - static_cast<u32>(__builtin_clz(q));
            const u32 sub = j >= split ? ((q - (1u << j)) >> (j - split)) : (q - (1u << j));
            return 1u + ((j << 8) | sub);
        }

        //! Rank of a request slot that is not in the group being served.
        constexpr u32 kNoRank = 0xffffffffu;

        //! The requests a claim serves: every lane holds M request slots, and
        //! rank[j] is slot j's rank among the requests still waiting
        //! (kNoRank, or any value at least the claim's count, for a slot
        //! that waits for nothing).  Claims hand out chunks in rank order,
        //! so a claim that served n chunks served ranks 0 .. n-1.
        template<u32 M>
        struct GroupRanks
        {
            u32 rank[M];
        };

        //! The waiting requests of a group of one request per lane: the
        //! lanes in `bits`, ranked by lane (uniform 64-bit state).
        struct LaneNeed
        {
            u64 bits;

            MALLOCMC_D bool open() const
            {
                return bits != 0u;
            }
            MALLOCMC_D u32 remaining() const
            {
                return static_cast<u32>(hw::popc(bits));
            }
            //! Progress marker: changes whenever a claim serves a request.
            MALLOCMC_D u32 mark() const
            {
                return remaining();
            }
            MALLOCMC_D GroupRanks<1> ranks() const
            {
                GroupRanks<1> r;
                r.rank[0] = ((bits >> hw::laneid()) & 1u) != 0u ? hw::rankBelow(bits) : kNoRank;
                return r;
            }
            //! The n lowest-ranked waiting requests were served.
            MALLOCMC_D void served(u32 n)
            {
                const GroupRanks<1> r = ranks();
                bits &= ~hw::ballot(r.rank[0] < n);
            }
        };

        //! The waiting requests of a group of M requests per lane: absolute
        //! ranks 0 .. K-1 (`rank`), of which the lowest `got` are served.
        template<u32 M>
        struct RankNeed
        {
            GroupRanks<M> rank;
            u32 K;
            u32 got;

            MALLOCMC_D bool open() const
            {
                return got < K;
            }
            MALLOCMC_D u32 remaining() const
            {
                return K - got;
            }
            MALLOCMC_D u32 mark() const
            {
                return got;
            }
            MALLOCMC_D GroupRanks<M> ranks() const
            {
                // below `got` and kNoRank wrap to values no claim reaches
                GroupRanks<M> r;
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    r.rank[j] = rank.rank[j] - got;
                return r;
            }
            MALLOCMC_D void served(u32 n)
            {
                got += n;
            }
        };

        //! Hands the chunks of `bits` (in bit order, chunk size cs, page data
        //! at `base`) to the group's ranks [from, from + n).
        //! `bits` is the same in every lane; when its set bits are the lowest
        //! ones (chunks 0 .. n-1 of a page claimed fresh), the r-th is bit r
        //! and the bit search is skipped.
        template<u32 M>
        MALLOCMC_D void handOut(const GroupRanks<M>& g, u32 from, u32 n, char* base, u32 bits, u32 cs,
                                char* (&out)[M])
        {
            if((bits & (bits + 1u)) == 0u)
            {
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    const u32 r = g.rank[j] - from; // wraps to a huge value below `from` and for kNoRank
                    if(r < n)
                        out[j] = base + static_cast<u64>(r) * cs;
                }
                return;
            }
#pragma unroll
            for(u32 j = 0; j < M; ++j)
            {
                const u32 r = g.rank[j] - from;
                if(r < n)
                    out[j] = base + static_cast<u64>(wave::selectBit(bits, r)) * cs;
            }
        }

        struct WaveAlloc
        {
            const ScatterHeap& h;
            WaveTuning t;

            //! Where the chunk bits of a page of chunk size cs live: the PTE
            //! bitmask (flat pages) or the page's segment masks (hierarchical).
            struct Layout
            {
                u32* words;
                u32 nWords;
                u32 full;
                u32 additional;
                bool hier;
            };

            MALLOCMC_D Layout layout(u32 page, u32 cs) const
            {
                Layout L;
                L.hier = cs <= h.hierarchyThreshold;
                if(L.hier)
                {
                    const u32 segmentsize = cs * 32u + 4u;
                    L.full = umin32(32u, h.pdiv(h.pagesize, segmentsize));
                    L.additional = h.calcAdditional(L.full, segmentsize, cs);
                    L.nWords = L.full + (L.additional > 0u ? 1u : 0u);
                    L.words = h.onPageMasks(page, L.nWords);
                }
                else
                {
                    L.full = 0u;
                    L.additional = 0u;
                    L.nWords = 1u;
                    L.words = h.pteBitmask(page);
                }
                return L;
            }

            //! pageCapacity(cs); the group's own chunk size c, whose capacity
            //! capC the walk computed once, skips the divisions (a divergent
            //! branch that no lane takes costs nothing).
            MALLOCMC_D u32 capacityOf(u32 cs, u32 c, u32 capC) const
            {
                u32 cap = capC;
                if(cs != c)
                    cap = h.pageCapacity(cs);
                return cap;
            }

            //! True for a chunk size the page reset applies to: chunked pages,
            //! and pagesize chunks with the pagesize-leak fix.  Pages of a
            //! multi-page allocation (chunk size above the page, count 0 while
            //! they are in use) are left alone.
            MALLOCMC_D bool resettable(u32 cs) const
            {
                const u32 P = h.pagesize;
                return cs != 0u && (cs < P || (cs == P && (h.flags & kFixPagesizeLeak)));
            }

            //! One lane: resets an empty chunked page (chunk size back to 0)
            //! if no one holds a chunk or a reservation on it.  The reference
            //! locks the count (0 -> pagesize), drops the chunk size and
            //! unlocks (Scatter.hpp:686-716); chunk size and count share one
            //! 64-bit word here, so a single CAS (cs, 0) -> (0, 0) does the
            //! same with no lock window.  No chunk bit is set at count zero,
            //! so no mask needs clearing.
            MALLOCMC_D void resetEmpty(u32 page) const
            {
                const u64 pr = atom::load(h.ptePair(page));
                const u32 cs = static_cast<u32>(pr);
                if((pr >> 32) != 0u || !resettable(cs))
                    return;
                atom::cas(h.ptePair(page), u64{cs}, u64{0});
            }

            //! One lane: returns `k` reserved slots of `page`.  With
            //! resetfreedpages, a return that leaves the count at zero means
            //! no chunk of the page is in use and no one else holds a
            //! reservation, so the page is reset as the last free would have
            //! (a reservation of a claim that backs off can hide the empty
            //! count from that free).
            MALLOCMC_D void backOff(u32 page, u32 k) const
            {
                const u32 old = atom::sub(h.pteCount(page), k);
                if(h.resetfreedpages && old == k)
                    resetEmpty(page);
            }

            //! Claims up to k chunks of size cs on `page` for the waiting ranks
            //! 0 .. k-1 of `g`.  Returns how many it served, lowest ranks
            //! first; each served slot's `out` gets its chunk.  `fresh`: the
            //! scan saw the page unformatted, so the claim formats it (CAS
            //! 0 -> cs).
            //!
            //! The reference's tryUsePage (Scatter.hpp:469-517) reserves with
            //! atomicAdd(count), then re-checks the chunk size with a CAS, then
            //! reads the mask word it will claim from.  Here one 64-bit add on
            //! the (chunk size, count) word reserves and returns the chunk size
            //! of the same instant, so the re-check needs no second round trip;
            //! the mask snapshot rides along (a hint only: the atomicOr's
            //! result decides which chunks are won).  A fresh page is formatted
            //! by a CAS issued just before the add on the same word.  The
            //! chunk size must be read by the reservation itself: a check made
            //! before it can be undone by a reset that completes before the
            //! reservation lands.  When the reservation finds the page
            //! unformatted, the claim formats it afterwards, as the reference
            //! does.
            template<u32 M>
            MALLOCMC_D u32 claimOnPage(u32 page, u32 cs, u32 bmax, u32 maxcs, u32 k, const GroupRanks<M>& g,
                                       bool fresh, char* (&out)[M]) const
            {
                const u32 me = hw::laneid();
                const u32 P = h.pagesize;
                bool formatted = false;
                u32 waited = 0u;
                for(;;)
                {
                    u32 old = 0u, rcs = 0u, fcs = 1u;
                    if(me == 0u)
                    {
                        if(fresh)
                            fcs = atom::cas(h.pteChunksize(page), 0u, cs);
                        const u64 pr = atom::add(h.ptePair(page), u64{k} << 32);
                        old = static_cast<u32>(pr >> 32);
                        rcs = static_cast<u32>(pr);
                    }
                    Layout L = layout(page, cs);
                    u32 snap = 0u;
                    if(me < L.nWords)
                        snap = atom::load(&L.words[me]);
                    old = wave::lane(old, 0u);
                    rcs = wave::lane(rcs, 0u);
                    if(wave::lane(fcs, 0u) == 0u)
                        formatted = true; // this wave (re)formatted the page
                    if(h.resetfreedpages && old >= P)
                    {
                        // page locked by a concurrent reset of the reference
                        // protocol (Scatter.hpp:686-716)
                        if(me == 0u)
                            backOff(page, k);
                        if(formatted && waited < kResetWaitLimit)
                        {
                            // D5 of scatter_heap.hpp, for the wave (bounded:
                            // after the limit the group moves on)
                            u32 spins = waited;
                            if(me == 0u)
                                while(atom::load(h.pteCount(page)) >= P && spins < kResetWaitLimit)
                                {
                                    __builtin_amdgcn_s_sleep(1);
                                    ++spins;
                                }
                            waited = wave::lane(spins, 0u) + 1u;
                            continue;
                        }
                        return 0u;
                    }
                    if(rcs == 0u)
                    {
                        // reserved on an unformatted page: format it under the
                        // reservation (no reset can run while the count is held)
                        u32 c2 = 0u;
                        if(me == 0u)
                            c2 = atom::cas(h.pteChunksize(page), 0u, cs);
                        c2 = wave::lane(c2, 0u);
                        if(c2 == 0u)
                            formatted = true;
                        rcs = c2 == 0u ? cs : c2;
                    }
                    bool snapValid = true;
                    if(rcs != cs)
                    {
                        if(rcs >= bmax && rcs <= maxcs)
                        {
                            cs = rcs; // formatted meanwhile with a fitting size
                            L = layout(page, cs);
                            snapValid = false;
                        }
                        else
                        {
                            if(me == 0u)
                                backOff(page, k);
                            return 0u;
                        }
                    }
                    const u32 cap = h.pageCapacity(cs);
                    const u32 grant = old < cap ? umin32(k, cap - old) : 0u;
                    const u32 got = grant != 0u ? claimBits<M>(page, cs, L, grant, g, out, snapValid, snap)
                                                : 0u;
                    if(got < k && me == 0u)
                    {
                        if(got == 0u)
                            backOff(page, k);
                        else
                            atom::sub(h.pteCount(page), k - got); // holds got chunks: count stays > 0
                    }
                    return got;
                }
            }

            //! Claims `grant` free bits on the page's mask words and hands the
            //! chunks to the waiting ranks of `g` in order.  Returns
            //! how many it handed out.  Pass 0 chooses from `snap0` (lane w:
            //! word w) when `haveSnap`.
            template<u32 M>
            MALLOCMC_D u32 claimBits(u32 page, u32 cs, const Layout& L, u32 grant, const GroupRanks<M>& g,
                                     char* (&out)[M], bool haveSnap, u32 snap0) const
            {
                const u32 me = hw::laneid();
                const u32 P = h.pagesize;
                const bool hier = L.hier;
                const u32 nWords = L.nWords, full = L.full, additional = L.additional;
                u32* const words = L.words;
                char* const data = h.pageData(page);
                u32 remaining = grant;
                u32 gotTotal = 0u;
                for(u32 pass = 0; pass < 3u && remaining > 0u; ++pass)
                {
                    // lane w < nWords snapshots mask word w
                    const bool owner = me < nWords;
                    u32 spots = 0u, snap = 0u;
                    if(owner)
                    {
                        spots = hier ? (me < full ? 32u : additional) : umin32(h.pdiv(P, cs), 32u);
                        snap = (pass == 0u && haveSnap) ? snap0 : atom::load(&words[me]);
                    }
                    const u32 sel = spots == 32u ? ~0u : ((1u << spots) - 1u);
                    const u32 freeBits = owner ? (~snap & sel) : 0u;
                    const u32 freeCnt = hw::popc(freeBits);
                    // exclusive prefix of free counts (<= 32) over words 0..me-1
                    const u32 prefix = wave::exclusiveSum(freeCnt);
                    u32 take = 0u;
                    if(owner && prefix < remaining)
                        take = umin32(freeCnt, remaining - prefix);
                    const u32 choose = wave::lowestBits(freeBits, take);
                    u32 gotBits = 0u;
                    if(choose != 0u)
                    {
                        const u32 old = atom::or_(&words[me], choose);
                        gotBits = choose & ~old;
                        // The reference also marks a segment it filled as full in
                        // the PTE bitmask (addChunkHierarchy, Scatter.hpp:408-419).
                        // The wave reads every mask word anyway, so it sets no
                        // such hint: a hint set after a concurrent free of that
                        // segment would outlive the page's chunks (a stale bit the
                        // next flat chunk size would read as a chunk in use).
                    }
                    // hand out the chunks word by word, in rank order
                    u64 words64 = hw::ballot(gotBits != 0u);
                    while(words64 != 0u)
                    {
                        const u32 w = hw::ffs(words64) - 1u;
                        words64 &= words64 - 1u;
                        const u32 bits = wave::lane(gotBits, w);
                        const u32 n = hw::popc(bits);
                        handOut<M>(g, gotTotal, n, data + static_cast<u64>(32u * w) * cs, bits, cs, out);
                        gotTotal += n;
                    }
                    remaining = grant - gotTotal;
                }
                return gotTotal;
            }

            //! Claims chunks for the waiting ranks 0 .. k-1 of `g` on several flat
            //! pages of one region at once.  Lane l stands for page `pageL` of the scan
            //! (chunk size pcs, count pcount, PTE bitmask pbits as the scan saw
            //! them); `used` / `fresh` are the candidate lanes.  Pages are
            //! taken in the scan's preference order (used pages from the
            //! rotation point, then fresh pages) until their free chunks cover
            //! the group.  Each chosen page's lane issues, in one round trip,
            //! the format CAS of a fresh page, the 64-bit count reservation
            //! (which returns the chunk size it was made under) and an
            //! atomicOr of the chunk bits it picked from the scanned bitmask:
            //! ownership of a chunk is decided by the atomicOr's result as in
            //! claimBits, so picking from a stale bitmask only costs chunks,
            //! never hands one out twice.  A page that turned out locked by a
            //! reset, or whose reservation saw another chunk size, gives back
            //! what it won.  Returns how many ranks it served, lowest first;
            //! `failed` gets the chosen pages that yielded no chunk (the scan's
            //! view of them was stale).
            template<u32 M>
            MALLOCMC_D u32 claimFlat(u32 pageL, u32 pcs, u32 pcount, u32 pbits, u64 used, u64 fresh, u32 rot,
                                     u32 c, u32 capC, u32 k, const GroupRanks<M>& g, u64& chosen, u64& failed,
                                     char* (&out)[M]) const
            {
                const u32 me = hw::laneid();
                const u32 P = h.pagesize;
                const bool isUsed = ((used >> me) & 1u) != 0u;
                const bool isFresh = ((fresh >> me) & 1u) != 0u;
                const u32 csl = isUsed ? pcs : c;
                const u32 capl = (isUsed || isFresh) ? capacityOf(csl, c, capC) : 0u;
                const u32 selBits = capl >= 32u ? ~0u : ((1u << capl) - 1u);
                const u32 freeBits = (isUsed || isFresh) ? (~pbits & selBits) : 0u;
                const u32 room = isUsed ? (capl > pcount ? capl - pcount : 0u) : capl;
                const u32 f = umin32(static_cast<u32>(hw::popc(freeBits)), room);
                // position of this page in the preference order
                const u64 hiMask = rot < 64u ? (~0ull << rot) : 0ull;
                const bool inUhi = isUsed && ((hiMask >> me) & 1u);
                const bool inUlo = isUsed && !inUhi;
                const bool inFhi = isFresh && ((hiMask >> me) & 1u);
                const bool inFlo = isFresh && !inFhi;
                const u32 pUhi = wave::exclusiveSum(inUhi ? f : 0u);
                const u32 pUlo = wave::exclusiveSum(inUlo ? f : 0u);
                const u32 pFhi = wave::exclusiveSum(inFhi ? f : 0u);
                const u32 pFlo = wave::exclusiveSum(inFlo ? f : 0u);
                const u32 tUhi = wave::lane(pUhi + (inUhi ? f : 0u), 63u);
                const u32 tUlo = wave::lane(pUlo + (inUlo ? f : 0u), 63u);
                const u32 tFhi = wave::lane(pFhi + (inFhi ? f : 0u), 63u);
                u32 pos = 0u;
                if(inUhi)
                    pos = pUhi;
                else if(inUlo)
                    pos = tUhi + pUlo;
                else if(inFhi)
                    pos = tUhi + tUlo + pFhi;
                else if(inFlo)
                    pos = tUhi + tUlo + tFhi + pFlo;
                const u32 take = (f != 0u && pos < k) ? umin32(f, k - pos) : 0u;
                const u32 choose = take != 0u ? wave::lowestBits(freeBits, take) : 0u;
                chosen = hw::ballot(take != 0u);
                // one round trip for every chosen page: a fresh page is
                // formatted first, then the 64-bit add reserves and returns the
                // chunk size of that instant (a reservation that finds the page
                // unformatted gives up: the format may have landed behind it)
                u32 old = 0u, rcs = 0u, fcs = 1u, oldbits = 0u;
                if(take != 0u)
                {
                    if(isFresh)
                        fcs = atom::cas(h.pteChunksize(pageL), 0u, csl);
                    const u64 pr = atom::add(h.ptePair(pageL), u64{take} << 32);
                    oldbits = atom::or_(h.pteBitmask(pageL), choose);
                    old = static_cast<u32>(pr >> 32);
                    rcs = static_cast<u32>(pr);
                }
                const bool formatted = take != 0u && fcs == 0u;
                const bool csOk = rcs == csl;
                const bool locked = h.resetfreedpages && old >= P;
                const u32 won = take != 0u ? (choose & ~oldbits) : 0u;
                const u32 keep = (take != 0u && csOk && !locked) ? won : 0u;
                const u32 kept = static_cast<u32>(hw::popc(keep));
                if(take != 0u)
                {
                    if(won & ~keep)
                        atom::and_(h.pteBitmask(pageL), ~(won & ~keep));
                    if(kept == 0u)
                    {
                        backOff(pageL, take);
                        if(formatted && locked)
                        {
                            // this lane formatted a page whose reset still held
                            // the count lock: once the reset is done, take the
                            // chunk size back unless someone uses the page (D5)
                            u32 spins = 0u;
                            while(atom::load(h.pteCount(pageL)) >= P && spins < kResetWaitLimit)
                            {
                                __builtin_amdgcn_s_sleep(1);
                                ++spins;
                            }
                            resetEmpty(pageL);
                        }
                    }
                    else if(kept < take)
                        atom::sub(h.pteCount(pageL), take - kept); // holds kept chunks: count stays > 0
                }
                failed = hw::ballot(take != 0u && kept == 0u);
                // hand the kept chunks to the group's ranks in order
                u32 got = 0u;
                for(u64 pagesKept = hw::ballot(kept != 0u); pagesKept != 0u; pagesKept &= pagesKept - 1u)
                {
                    const u32 l = hw::ffs(pagesKept) - 1u;
                    const u32 bits = wave::lane(keep, l);
                    const u32 n = static_cast<u32>(hw::popc(bits));
                    handOut<M>(g, got, n, h.pageData(wave::lane(pageL, l)), bits, wave::lane(csl, l), out);
                    got += n;
                }
                return got;
            }

            //! Claims chunks of a hierarchical chunk size c for the waiting
            //! ranks 0 .. k-1 of `g` on several pages the scan saw unformatted, in
            //! one round trip.  Lane l stands for page `pageL` (candidates in
            //! `fresh`); pages are taken from the rotation point on until their
            //! capacity covers the group.  Each chosen page's lane formats its
            //! page (CAS 0 -> c), reserves `take` slots with the 64-bit add
            //! (which returns the chunk size it was made under) and sets the
            //! lowest `take` chunk bits in the page's first segment words, all
            //! at once: a page gives a group at most 32 NW chunks (NW = 2 words
            //! for one request per lane, 4 for more), or its whole capacity
            //! when that is smaller.  The reference's
            //! tryUsePage (Scatter.hpp:469-517) reserves, then re-checks the
            //! chunk size, then claims a bit, one round trip each; claimOnPage
            //! needs two.  Ownership of a chunk is decided by the atomicOr's
            //! result, as in claimBits: a concurrent claim of the same page only
            //! costs chunks.  A page that turned out formatted with another
            //! size, or locked by a reset, gives back what it won.  Returns how
            //! many ranks it served, lowest first; `failed` gets the chosen
            //! pages that yielded no chunk.
            template<u32 M>
            MALLOCMC_D u32 claimFresh(u32 pageL, u64 fresh, u32 rot, u32 c, u32 capC, const Layout& LC, u32 k,
                                      const GroupRanks<M>& g, u64& chosen, u64& failed, char* (&out)[M]) const
            {
                constexpr u32 NW = M >= 2u ? 4u : 2u; // segment words one page's claim may touch
                const u32 P = h.pagesize;
                const bool isFresh = ((fresh >> hw::laneid()) & 1u) != 0u;
                const u32 cap = capC;
                // order among the candidates: from the rotation point, then the rest
                const u64 hiMask = rot < 64u ? (~0ull << rot) : 0ull;
                const u64 below = hw::lanemask_lt();
                const u32 nHi = static_cast<u32>(hw::popc(fresh & hiMask));
                const bool inHi = isFresh && ((hiMask >> hw::laneid()) & 1u) != 0u;
                const u32 order = inHi ? static_cast<u32>(hw::popc(fresh & hiMask & below))
                                       : nHi + static_cast<u32>(hw::popc(fresh & ~hiMask & below));
                const u32 perPage = umin32(cap, 32u * NW);
                const u32 pos = order * perPage;
                const u32 take = (isFresh && pos < k) ? umin32(perPage, k - pos) : 0u;
                Layout L = LC; // the chunk size's layout; this lane's page's words
                L.words = h.onPageMasks(pageL, LC.nWords);
                // the lowest `take` chunks: word w holds 32 (w < full) or
                // `additional` (w == full) of them
                u32 m[NW];
                u32 left = take;
#pragma unroll
                for(u32 w = 0; w < NW; ++w)
                {
                    const u32 sp = w < L.full ? 32u : (w == L.full ? L.additional : 0u);
                    const u32 tw = umin32(left, sp);
                    m[w] = tw >= 32u ? ~0u : ((1u << tw) - 1u);
                    left -= tw;
                }
                chosen = hw::ballot(take != 0u);
                u32 old = 0u, rcs = 0u, fcs = 1u;
                u32 ob[NW];
#pragma unroll
                for(u32 w = 0; w < NW; ++w)
                    ob[w] = 0u;
                if(take != 0u)
                {
                    fcs = atom::cas(h.pteChunksize(pageL), 0u, c);
                    const u64 pr = atom::add(h.ptePair(pageL), u64{take} << 32);
                    ob[0] = atom::or_(&L.words[0], m[0]);
#pragma unroll
                    for(u32 w = 1; w < NW; ++w)
                        if(m[w] != 0u)
                            ob[w] = atom::or_(&L.words[w], m[w]);
                    old = static_cast<u32>(pr >> 32);
                    rcs = static_cast<u32>(pr);
                }
                const bool formatted = take != 0u && fcs == 0u;
                const bool locked = h.resetfreedpages && old >= P;
                const bool usable = take != 0u && rcs == c && !locked;
                u32 keep[NW];
                u32 kept = 0u;
#pragma unroll
                for(u32 w = 0; w < NW; ++w)
                {
                    const u32 won = take != 0u ? (m[w] & ~ob[w]) : 0u;
                    keep[w] = usable ? won : 0u;
                    kept += static_cast<u32>(hw::popc(keep[w]));
                    if(take != 0u && (won & ~keep[w]) != 0u)
                        atom::and_(&L.words[w], ~(won & ~keep[w]));
                }
                if(take != 0u)
                {
                    if(kept == 0u)
                    {
                        backOff(pageL, take);
                        if(formatted && locked)
                        {
                            // formatted a page whose reset still held the count
                            // lock: once the reset is done, take the chunk size
                            // back unless someone uses the page (D5)
                            u32 spins = 0u;
                            while(atom::load(h.pteCount(pageL)) >= P && spins < kResetWaitLimit)
                            {
                                __builtin_amdgcn_s_sleep(1);
                                ++spins;
                            }
                            resetEmpty(pageL);
                        }
                    }
                    else if(kept < take)
                        atom::sub(h.pteCount(pageL), take - kept); // holds kept chunks: count stays > 0
                }
                failed = hw::ballot(take != 0u && kept == 0u);
                // hand the kept chunks to the group's ranks in order, word by word
                u32 got = 0u;
                for(u64 pagesKept = hw::ballot(kept != 0u); pagesKept != 0u; pagesKept &= pagesKept - 1u)
                {
                    const u32 l = hw::ffs(pagesKept) - 1u;
                    char* const data = h.pageData(wave::lane(pageL, l));
#pragma unroll
                    for(u32 w = 0; w < NW; ++w)
                    {
                        const u32 bits = wave::lane(keep[w], l);
                        const u32 n = static_cast<u32>(hw::popc(bits));
                        handOut<M>(g, got, n, data + static_cast<u64>(32u * w) * c, bits, c, out);
                        got += n;
                    }
                }
                return got;
            }

            //! Skips ahead over regions without a candidate page for chunk size
            //! c, starting `from` regions after region0 and stopping at the
            //! first region that has one or at `numregions`.  One round trip
            //! reads the fill levels of the next 64 regions (lane r: region r)
            //! and the PTEs of the first 256 pages (each lane reads 4 pages).
            //! Regions filled past `checklevel` are passed over without looking
            //! at their pages, as the walk does (Scatter.hpp:573-575), so a
            //! stretch of such regions costs one round trip per 64 regions;
            //! open regions are examined page by page, 256 pages per round
            //! trip.  Open regions passed over get the fill-level increment an
            //! unsuccessful visit gives them (Scatter.hpp:612-617; the CAS
            //! starts from the level read after the visit: from the level read
            //! before it, concurrent visits cancel each other's increments and
            //! measured slower), and the access-block cursor advances as the
            //! walk would have (:620-629).  Returns the regions skipped.  For
            //! regionsize < 64 only.
            template<class Wrap>
            MALLOCMC_D u32 fastForward(u32 region0, u32 from, u32 numregions, u32 checklevel, u32 bmax, u32 maxcs,
                                       u32 c, u32 capC, Wrap wrap) const
            {
                const u32 me = hw::laneid();
                const u32 R = h.regionsize;
                const u32 per = umin32(64u, h.divR.div(256u)); // regions whose pages one round trip reads
                u32 skipped = 0u;
                while(from + skipped < numregions)
                {
                    const u32 nregs = umin32(64u, numregions - from - skipped);
                    const u32 npage = umin32(per, nregs);
                    // lane r: fill level of region r of this round
                    u32 fill = 0xffffffffu;
                    if(me < nregs)
                        fill = atom::load(&h.regions[wrap(region0 + from + skipped + me)]);
                    u32 css[4], cnts[4];
                    u32 regRel[4];
#pragma unroll
                    for(u32 l = 0; l < 4u; ++l)
                    {
                        const u32 pageRel = l * 64u + me;
                        regRel[l] = h.divR.div(pageRel);
                        css[l] = 0u;
                        cnts[l] = 0u;
                        if(regRel[l] < npage)
                        {
                            const u32 region = wrap(region0 + from + skipped + regRel[l]);
                            const u32 page = region * R + (pageRel - regRel[l] * R);
                            // chunk size and count in one load of their paired word
                            const u64 pr = atom::load(h.ptePair(page));
                            css[l] = static_cast<u32>(pr);
                            cnts[l] = static_cast<u32>(pr >> 32);
                        }
                    }
                    const u64 openMask = hw::ballot(me < nregs && fill < checklevel);
                    u32 first = 0xffffffffu;
#pragma unroll
                    for(u32 l = 0; l < 4u; ++l)
                    {
                        if(regRel[l] >= npage || ((openMask >> regRel[l]) & 1u) == 0u)
                            continue;
                        const u32 pc = css[l];
                        const bool cand = pc == 0u
                            || (pc >= bmax && pc <= maxcs && pc <= h.pagesize && cnts[l] < capacityOf(pc, c, capC));
                        if(cand)
                            first = umin32(first, regRel[l]);
                    }
                    first = wave::min(first);
                    u32 passed;
                    if(first != 0xffffffffu)
                        passed = first;
                    else
                    {
                        // past the examined regions, up to the next open one
                        const u64 later = npage < 64u ? (openMask >> npage) << npage : 0u;
                        passed = later != 0u ? hw::ffs(later) - 1u : nregs;
                    }
                    // open regions whose pages were examined without success
                    const u32 examined = umin32(passed, npage);
                    if(me < examined && fill < checklevel)
                    {
                        u32* level = &h.regions[wrap(region0 + from + skipped + me)];
                        const u32 f = atom::load(level);
                        if(f < checklevel && f + 1u <= R)
                            atom::cas(level, f, f + 1u);
                    }
                    // access-block cursor over the regions passed (none to move
                    // on a heap of one access block)
                    if(h.accessblocks > 1u)
                    {
                        u32 abid = 0u;
                        if(me < passed)
                        {
                            const u32 rg = wrap(region0 + from + skipped + me + 1u);
                            const u32 q = h.divPpb.div(rg * R);
                            if(q * h.pagesperblock == rg * R)
                                abid = q;
                        }
                        abid = wave::max(abid);
                        if(abid != 0u && me == 0u && abid > atom::load(&h.scalars->firstfreeblock))
                            atom::store(&h.scalars->firstfreeblock, abid);
                    }
                    skipped += passed;
                    if(first != 0xffffffffu)
                        break;
                }
                return skipped;
            }

            //! Group allocation of chunk size c for the lanes in `group`.  Walks
            //! the regions in the reference's order (Scatter.hpp:563-635: from the
            //! hashed page, region by region, two passes with the 3/4 fill
            //! threshold first) but examines up to 64 pages -- 64/regionsize
            //! regions -- per round trip.
            //! `ffb`: the access-block cursor (_firstfreeblock) as the wave read it.
            //! `need` holds the group's waiting requests (LaneNeed: one per
            //! lane; RankNeed: M per lane); the returned copy records which
            //! were served, each with its chunk in `out`.  (Passed by value:
            //! by reference, the compiler kept the wave malloc's walk in more
            //! vector registers, 128 instead of 76.)
            template<u32 M, class Need>
            MALLOCMC_D Need allocGroup(Need need, u32 c, u32 bmax, u32 maxcs, u32 ffb, char* (&out)[M]) const
            {
                const u32 me = hw::laneid();
                const u32 R = h.regionsize;
                const u32 ppb = h.pagesperblock;
                const u32 numregions = h.numregions;
                // the wave term of Scatter.hpp:535-538; the wave's id within its CU
                // (slot and SIMD) rather than the slot alone, which the same-slot
                // waves of the CU's four SIMDs share: with it, the waves of a
                // 1,024-thread block that serve one size class started their
                // walks on the same page and raced for the same pages
                const u32 hashBase = c * h.hashingK + h.hashingDistMP * hw::cuid()
                    + (h.hashingDistWP + h.hashingDistWPRel * h.divP.div(64u * c)) * hw::waveInCu();
                u32 startInBlock;
                if(t.xcdSlices > 1u && ppb >= t.xcdSlices * R)
                {
                    if(t.xcdSlices == 8u)
                        startInBlock = (hw::xccid() & 7u) * h.divSlice8.d + h.divSlice8.mod(hashBase);
                    else
                    {
                        const u32 slice = ppb / t.xcdSlices;
                        startInBlock = (hw::xccid() % t.xcdSlices) * slice + hashBase % slice;
                    }
                }
                else
                    startInBlock = h.divPpb.mod(hashBase);
                const u32 globalStart = startInBlock + ffb * ppb;
                const u32 region0 = h.divR.div(globalStart);
                const u32 off0 = globalStart - region0 * R; // every region is entered at this page
                // region0 + x for 0 <= x < 2 * numregions, wrapped
                auto wrap = [numregions](u32 r) {
                    r = r >= numregions ? r - numregions : r;
                    return r >= numregions ? r - numregions : r;
                };
                // regions examined per round trip
                const u32 W = R >= 64u ? 1u : h.divR.div(64u);
                const u32 span = R >= 64u ? 64u : W * R;
                u32 checklevel = R * 3u / 4u;
                // chunks above the hierarchy threshold: every fitting page keeps
                // its chunk bits in the PTE bitmask (candidates have a chunk
                // size >= bmax = c)
                const bool flat = c > h.hierarchyThreshold;
                // claims on several pages per round trip (claimFlat, claimFresh);
                // with flat_single every page is claimed on its own (claimOnPage)
                const bool multi = t.flatFast != 0u;
                // the group's chunk size: capacity and segment layout, once
                const u32 capC = h.pageCapacity(c);
                const Layout LC = layout(0u, c);
                for(u32 finder = 0; finder < 2u && need.open(); ++finder)
                {
                    for(u32 done = 0; done < numregions && need.open(); done += W)
                    {
                        const u32 nreg = umin32(W, numregions - done);
                        bool windowHadCand = false;
                        // lane -> (region of this round, page within region)
                        const u32 ri = R >= 64u ? 0u : h.divR.div(me);
                        const u32 region = wrap(region0 + done + ri);
                        const bool inSpan = me < span && ri < nreg;
                        for(u32 base = 0; base < (R >= 64u ? R : 1u) && need.open(); base += 64u)
                        {
                            // Claims race with the other waves that scanned the same
                            // pages: a round whose pages were taken meanwhile stops and
                            // the window is read again, instead of trying the rest of a
                            // stale snapshot page by page (at most kWindowReads reads).
                            u64 raised = 0u; // regions of the window whose fill level this visit raised
                            for(u32 read = 0; read < kWindowReads && need.open(); ++read)
                            {
                                bool stale = false;
                                u32 fill = 0xffffffffu, pcs = 0u, pcount = 0u, pbits = 0u;
                                const u32 pj = R >= 64u ? base + me : me - ri * R;
                                const bool valid = inSpan && pj < R;
                                if(valid)
                                {
                                    fill = atom::load(&h.regions[region]);
                                    const u64 pr = atom::load(h.ptePair(region * R + pj));
                                    pcs = static_cast<u32>(pr);
                                    pcount = static_cast<u32>(pr >> 32);
                                    if(flat && multi)
                                        pbits = atom::load(h.pteBitmask(region * R + pj));
                                }
                                const bool regionOpen = valid && fill < checklevel;
                                bool cand = false;
                                if(regionOpen)
                                {
                                    if(pcs == 0u)
                                        cand = true;
                                    else if(pcs >= bmax && pcs <= maxcs && pcs <= h.pagesize)
                                        cand = pcount < capacityOf(pcs, c, capC);
                                }
                                const u64 usedAll = hw::ballot(cand && pcs != 0u);
                                const u64 freeAll = hw::ballot(cand && pcs == 0u);
                                const u64 open = hw::ballot(regionOpen);
                                windowHadCand = windowHadCand || (usedAll | freeAll) != 0u;
                                const u32 rounds = R >= 64u ? 1u : nreg;
                                for(u32 k = 0; k < rounds && need.open(); ++k)
                                {
                                    const u32 lo = R >= 64u ? 0u : k * R;
                                    const u32 width = R >= 64u ? umin32(64u, R - base) : R;
                                    const u64 sel = (width >= 64u ? ~0ull : ((1ull << width) - 1u)) << lo;
                                    if((open & sel) == 0u)
                                        continue; // region filled past the threshold
                                    const u32 fillK = wave::lane(fill, lo);
                                    const u32 regionK = R >= 64u ? region : wrap(region0 + done + k);
                                    u64 used = usedAll & sel;
                                    u64 cands = freeAll & sel;
                                    const u32 rotAbs = R >= 64u ? (off0 >= base ? off0 - base : 0u) : off0;
                                    const u32 rot = lo + rotAbs;
                                    const u32 before = need.mark();
                                    bool staleK = false;
                                    if(multi)
                                    {
                                        const u32 pageL = regionK * R + (R >= 64u ? base + me : me - lo);
                                        if(flat)
                                        {
                                            // rounds over the scanned candidates, used pages
                                            // first; pages a round chose are not offered again
                                            while(need.open() && (used | cands) != 0u)
                                            {
                                                u64 chosen = 0u, failed = 0u;
                                                need.served(wave::uniform(claimFlat<M>(pageL, pcs, pcount, pbits, used, cands, rot, c, capC,
                                                                         need.remaining(), need.ranks(), chosen,
                                                                         failed, out)));
                                                used &= ~chosen;
                                                cands &= ~chosen;
                                                staleK = staleK || failed != 0u;
                                                if(chosen == 0u || failed != 0u)
                                                    break;
                                            }
                                        }
                                        else
                                        {
                                            // partially used pages of a fitting chunk size first,
                                            // one claim each: formatting a fresh page while a
                                            // fitting one has room is what fragments the
                                            // reference's heap; then the unformatted pages,
                                            // several per round trip
                                            while(used != 0u && need.open())
                                            {
                                                const u64 hi = rot < 64u ? (used >> rot) << rot : 0u;
                                                const u32 l = hw::ffs(hi != 0u ? hi : used) - 1u;
                                                used &= ~(1ull << l);
                                                const u32 page = regionK * R + (R >= 64u ? base + l : l - lo);
                                                const u32 served = wave::uniform(claimOnPage<M>(page, wave::lane(pcs, l), bmax,
                                                                                  maxcs, need.remaining(), need.ranks(),
                                                                                  false, out));
                                                need.served(served);
                                                if(served == 0u)
                                                {
                                                    staleK = true;
                                                    break;
                                                }
                                            }
                                            while(t.freshFast != 0u && need.open() && cands != 0u && !staleK)
                                            {
                                                u64 chosen = 0u, failed = 0u;
                                                need.served(wave::uniform(claimFresh<M>(pageL, cands, rot, c, capC, LC, need.remaining(),
                                                                          need.ranks(), chosen, failed, out)));
                                                cands &= ~chosen;
                                                staleK = staleK || failed != 0u;
                                                if(chosen == 0u || failed != 0u)
                                                    break;
                                            }
                                        }
                                        used = 0u;
                                        if(flat || t.freshFast != 0u)
                                            cands = 0u;
                                    }
                                    // one page at a time (flat_single, and the unformatted pages
                                    // of fresh_single), partially used pages of a fitting chunk
                                    // size first, free pages after them
                                    while((used | cands) != 0u && need.open())
                                    {
                                        const u64 set = used != 0u ? used : cands;
                                        const u64 hi = rot < 64u ? (set >> rot) << rot : 0u;
                                        const u64 pick = hi != 0u ? hi : set;
                                        const u32 l = hw::ffs(pick) - 1u;
                                        used &= ~(1ull << l);
                                        cands &= ~(1ull << l);
                                        const u32 page = regionK * R + (R >= 64u ? base + l : l - lo);
                                        const u32 pc = wave::lane(pcs, l);
                                        // a free page is formatted by the claim itself
                                        need.served(wave::uniform(claimOnPage<M>(page, pc == 0u ? c : pc, bmax, maxcs,
                                                                   need.remaining(), need.ranks(), pc == 0u, out)));
                                    }
                                    stale = stale || staleK;
                                    // an unsuccessful visit raises the region's fill level
                                    // once (Scatter.hpp:612-617); a visit cut short by a
                                    // stale snapshot is not one yet
                                    if(need.mark() == before && !staleK)
                                    {
                                        if(((raised >> k) & 1u) == 0u && fillK + 1u <= R && me == 0u)
                                            atom::cas(&h.regions[regionK], fillK, fillK + 1u);
                                        raised |= 1ull << k;
                                    }
                                }
                                if(!stale)
                                    break;
                            }
                        }
                        // access-block cursor: advanced when the walk enters a new block
                        // (Scatter.hpp:620-629); lane i looks at the region after region i
                        // of this round (a heap of one access block has no other block)
                        if(h.accessblocks > 1u)
                        {
                            u32 abid = 0u;
                            if(me < nreg)
                            {
                                const u32 rg = wrap(region0 + done + me + 1u);
                                const u32 q = h.divPpb.div(rg * R);
                                if(q * ppb == rg * R)
                                    abid = q;
                            }
                            abid = wave::max(abid);
                            if(abid != 0u && me == 0u && abid > atom::load(&h.scalars->firstfreeblock))
                                atom::store(&h.scalars->firstfreeblock, abid);
                        }
                        // a window with no candidate page at all: skip ahead over
                        // the regions that have none either, 256 pages per round trip
                        if(!windowHadCand && need.open() && R < 64u && t.fastForward != 0u && done + W < numregions)
                            done += fastForward(region0, done + W, numregions, checklevel, bmax, maxcs, c, capC, wrap);
                    }
                    checklevel = R + 1u;
                }
                return need;
            }

            //! Marks pages [sp, sp + k) as one run of `bytes` (chunk size =
            //! bytes on every page), 64 pages per round trip: lane l CASes
            //! pages sp + l, sp + l + 64, ... from 0 to bytes in order and
            //! stops at its first page already taken; if any lane met one,
            //! every lane sets the pages it marked back to 0.  Only its own:
            //! a page that another claim of the same `bytes` holds reads the
            //! same value, and resetting it would free a run that claim won.
            //! The end state is that of the reference's markpages
            //! (Scatter.hpp:745-764), which CASes the pages one by one and
            //! rolls back the ones before the first taken page.
            MALLOCMC_D bool markpagesWave(u32 sp, u32 k, u32 bytes) const
            {
                const u32 me = hw::laneid();
                u32 marked = 0u; // this lane's pages sp + me + 64 * x, x < marked, hold bytes
                bool lost = false;
                for(u32 i = me; i < k; i += 64u)
                {
                    if(atom::cas(h.pteChunksize(sp + i), 0u, bytes) != 0u)
                    {
                        lost = true;
                        break;
                    }
                    ++marked;
                }
                if(hw::ballot(lost) == 0u)
                    return true;
                for(u32 x = 0; x < marked; ++x)
                    atom::cas(h.pteChunksize(sp + me + 64u * x), bytes, 0u);
                return false;
            }

            //! A run of k > 32 pages (the reference serializes every run on
            //! one mutex, Scatter.hpp:824-840).  Same top-down order from
            //! _firstFreePageBased as allocRun, one 64-page window per round
            //! trip without overlap: lane i learns how many free pages start
            //! at its page (the free pages above it in the window, plus the
            //! free run the window above ended with when its own run reaches
            //! the window's top), and the highest page with k or more is
            //! claimed with markpagesWave.  The cursor is not moved for such
            //! runs: the per-lane path scans page by page from it and needs
            //! every page above it in use, and a run across windows leaves no
            //! cheap proof that no shorter run lies above it.
            MALLOCMC_D char* allocLongRun(u32 bytes, u32 k) const
            {
                const u32 me = hw::laneid();
                const u32 numpages = h.numpages;
                if(k > numpages)
                    return nullptr;
                u32 start = 0u;
                if(me == 0u)
                    start = atom::load(&h.scalars->firstFreePageBased);
                start = wave::lane(start, 0u);
                if(start >= numpages)
                    start = numpages - 1u;
                for(u32 sweep = 0; sweep < 2u; ++sweep)
                {
                    long long top = sweep == 0u ? static_cast<long long>(start) : static_cast<long long>(numpages) - 1;
                    const long long bottom = sweep == 0u ? 0 : static_cast<long long>(start);
                    // free pages starting at top + 1 and going up (only what this
                    // sweep has seen: the first window of a sweep starts with 0)
                    u32 runAbove = 0u;
                    while(top >= bottom)
                    {
                        const long long base = top - 63 > bottom ? top - 63 : bottom;
                        const u32 width = static_cast<u32>(top - base) + 1u;
                        const long long pg = base + me;
                        bool isFree = false;
                        if(me < width)
                            isFree = atom::load(h.pteChunksize(static_cast<u32>(pg))) == 0u;
                        const u64 m = hw::ballot(isFree);
                        // free pages from this lane's page up to the window's top
                        const u64 upFromMe = me < width ? (m >> me) : 0u;
                        const u32 ones = upFromMe == ~0ull ? 64u : static_cast<u32>(__builtin_ctzll(~upFromMe));
                        const u32 len = isFree ? (ones >= width - me ? ones + runAbove : ones) : 0u;
                        u64 cand = hw::ballot(isFree && len >= k);
                        while(cand != 0u)
                        {
                            const u32 pos = 63u - static_cast<u32>(__builtin_clzll(cand));
                            const u32 sp = static_cast<u32>(base) + pos;
                            if(sp + k <= numpages && markpagesWave(sp, k, bytes))
                                return h.pageData(sp);
                            cand &= ~(1ull << pos);
                        }
                        runAbove = wave::lane(len, 0u); // the free run that starts at base
                        if(base == bottom)
                            break;
                        top = base - 1;
                    }
                }
                return nullptr;
            }

            //! Multi-page request of `bytes` > pagesize for one lane, served by the
            //! whole wave (uniform `bytes`).  The reference takes a global spin
            //! mutex and scans one PTE at a time from _firstFreePageBased down
            //! (Scatter.hpp:775-840); here the wave reads 64 PTEs per round trip
            //! in the same top-down order, finds runs of free pages with ballot
            //! arithmetic and claims one with markpages, whose CAS-and-roll-back
            //! (Scatter.hpp:745-764) is safe without the mutex.
            MALLOCMC_D char* allocRun(u32 bytes) const
            {
                const u32 me = hw::laneid();
                const u32 P = h.pagesize;
                const u32 numpages = h.numpages;
                const u32 k = divup(bytes, P);
                if(k > 32u)
                    return allocLongRun(bytes, k);
                u32 start = 0u;
                if(me == 0u)
                    start = atom::load(&h.scalars->firstFreePageBased);
                start = wave::lane(start, 0u);
                if(start >= numpages)
                    start = numpages - 1u;
                for(u32 sweep = 0; sweep < 2u; ++sweep)
                {
                    long long top = sweep == 0u ? static_cast<long long>(start) : static_cast<long long>(numpages) - 1;
                    const long long bottom = sweep == 0u ? 0 : static_cast<long long>(start);
                    // no two adjacent free pages seen between the cursor and the
                    // windows so far (a run needs two pages or more, so a single
                    // free page between runs cannot serve one)
                    bool usedAbove = sweep == 0u;
                    while(top >= bottom)
                    {
                        const long long base = top - 63 > bottom ? top - 63 : bottom;
                        const long long pg = base + me;
                        bool isFree = false;
                        if(pg <= top)
                            isFree = atom::load(h.pteChunksize(static_cast<u32>(pg))) == 0u;
                        const u64 m = hw::ballot(isFree);
                        const u64 pairs = m & (m >> 1); // bit i: pages i and i+1 free
                        u64 r = m;
                        for(u32 j = 1; j < k; ++j)
                            r &= m >> j;
                        while(r != 0u)
                        {
                            const u32 pos = 63u - static_cast<u32>(__builtin_clzll(r));
                            const u32 sp = static_cast<u32>(base) + pos;
                            u32 ok = 0u;
                            if(me == 0u)
                                ok = h.markpages(sp, k, bytes) ? 1u : 0u;
                            ok = wave::lane(ok, 0u);
                            if(ok)
                            {
                                // no run could start between the cursor and this
                                // run: move the cursor below the run, as the
                                // reference does when every page above was in use
                                // (Scatter.hpp:791-798).  Runs claimed concurrently
                                // leave the cursor behind the pages they took, so
                                // lower it with a min instead of the reference's
                                // CAS from the value read.  The per-lane path
                                // relies on this: it scans page by page from the
                                // cursor down, so no free run may lie above it.
                                const u64 above = pos + k < 64u ? (pairs >> (pos + k)) : 0u;
                                if(me == 0u && usedAbove && above == 0u && sp > 0u)
                                    atom::min(&h.scalars->firstFreePageBased, sp - 1u);
                                return h.pageData(sp);
                            }
                            r &= ~(1ull << pos);
                        }
                        // free pairs of this window lie above any run found lower
                        if(pairs != 0u)
                            usedAbove = false;
                        if(base == bottom)
                            break;
                        top = base + static_cast<long long>(k) - 2; // windows overlap by k-1 pages
                    }
                }
                return nullptr;
            }

            //! Wave malloc: every lane passes its (unpadded) request size.
            template<class Dist>
            MALLOCMC_D void* malloc(u32 bytesIn) const
            {
                const u32 me = hw::laneid();
                // the access-block cursor, read once for all groups of the wave
                // (the reference reads it per request; any recent value is a
                // valid start, Scatter.hpp:557-560); issued first so that its
                // round trip overlaps the size loads.  A heap of one access
                // block keeps it at 0 (no walk enters another block), and its
                // first scan window need not wait for it.
                u32 ffb = 0u;
                if(me == 0u && h.accessblocks > 1u)
                    ffb = atom::load(&h.scalars->firstfreeblock);
                Dist dist;
                const u32 b = dist.collect(h, h.pad(bytesIn));
                char* result = nullptr;
                bool want = b != 0u;
                // requests above a page: one wave-wide run claim each.  Once a
                // run of k pages was not found anywhere, the wave's later
                // requests of k or more pages get null without another search
                // (as if served at the same instant).
                u32 failedPages = 0xffffffffu;
                for(u64 big = hw::ballot(want && b > h.pagesize); big != 0u; big &= big - 1u)
                {
                    const u32 l = hw::ffs(big) - 1u;
                    const u32 bl = wave::lane(b, l);
                    const u32 kl = h.divP.div(bl + h.pagesize - 1u);
                    char* r = nullptr;
                    if(kl < failedPages)
                    {
                        r = allocRun(bl);
                        if(r == nullptr)
                            failedPages = kl;
                    }
                    if(me == l)
                        result = r;
                }
                if(b > h.pagesize)
                    want = false;
                u64 pending = hw::ballot(want);
                const u32 cls = want ? sizeClass(h, b, t.classSplit) : 0xffffffffu;
                ffb = wave::lane(ffb, 0u);
                while(pending != 0u)
                {
                    const u32 leader = hw::ffs(pending) - 1u;
                    const u32 lcls = wave::lane(cls, leader);
                    const bool inGroup = ((pending >> me) & 1u) != 0u && cls == lcls;
                    u64 group = hw::ballot(inGroup);
                    // cap the group at the aggregation width
                    if(t.aggregation < 64u && hw::popc(group) > t.aggregation)
                    {
                        // the lowest `aggregation` lanes of the group
                        const u32 lo = static_cast<u32>(group), hi = static_cast<u32>(group >> 32);
                        const u32 nlo = static_cast<u32>(hw::popc(lo));
                        group = t.aggregation <= nlo
                            ? static_cast<u64>(wave::lowestBits(lo, t.aggregation))
                            : (static_cast<u64>(lo)
                               | (static_cast<u64>(wave::lowestBits(hi, t.aggregation - nlo)) << 32));
                    }
                    const bool member = ((group >> me) & 1u) != 0u;
                    const u32 bmax = wave::max(member ? b : 0u);
                    const u32 c = umax32(bmax, h.paddedMinChunk); // the largest member's chunk
                    const u32 bmin = wave::min(member ? b : 0xffffffffu);
                    const u32 maxcs = umin32(h.pagesize, umax32(h.wastefactor * bmin, h.paddedMinChunk));
                    char* mine[1] = {nullptr};
                    const LaneNeed need = allocGroup<1>(LaneNeed{group}, c, bmax, maxcs, ffb, mine);
                    if(member && ((need.bits >> me) & 1u) == 0u)
                        result = mine[0];
                    // lanes of the group not served: the heap has no room for them
                    pending &= ~group;
                }
                void* r = result;
                if(isOOM(r, b))
                    r = nullptr;
                return dist.distribute(r);
            }

            //! One request set of mallocMany: pads the sizes, serves the
            //! requests above a page (one wave-wide run claim each, as in
            //! malloc(); a run of k pages that was not found ends the search
            //! for runs of k or more, `failedPages`), and returns the other
            //! requests as pending, with their size classes.
            template<u32 M>
            MALLOCMC_D void beginSet(const u32 (&bytesIn)[M], u32 (&b)[M], char* (&result)[M], u64 (&pending)[M],
                                     u32 (&cls)[M], u32& failedPages) const
            {
                const u32 me = hw::laneid();
                bool want[M];
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    b[j] = h.pad(bytesIn[j]);
                    result[j] = nullptr;
                    want[j] = b[j] != 0u;
                }
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    for(u64 big = hw::ballot(want[j] && b[j] > h.pagesize); big != 0u; big &= big - 1u)
                    {
                        const u32 l = hw::ffs(big) - 1u;
                        const u32 bl = wave::lane(b[j], l);
                        const u32 kl = h.divP.div(bl + h.pagesize - 1u);
                        char* r = nullptr;
                        if(kl < failedPages)
                        {
                            r = allocRun(bl);
                            if(r == nullptr)
                                failedPages = kl;
                        }
                        if(me == l)
                            result[j] = r;
                    }
                    if(b[j] > h.pagesize)
                        want[j] = false;
                }
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    pending[j] = hw::ballot(want[j]);
                    cls[j] = want[j] ? sizeClass(h, b[j], t.classSplit) : 0xffffffffu;
                }
            }

            //! Serves one size-class group of the pending requests: the class
            //! of the lowest pending request in (slot, lane) order (slot lj),
            //! at most `cap` of its requests in that order.  Chunks go straight
            //! to `result` (only served ranks are written); the group's
            //! requests leave `pending`, served or not (not served: the heap
            //! has no room for them).
            template<u32 M>
            MALLOCMC_D void serveGroup(u64 (&pending)[M], const u32 (&cls)[M], const u32 (&b)[M], char* (&result)[M],
                                       u32 lj, u32 cap, u32 ffb) const
            {
                const u32 me = hw::laneid();
                u64 lpend = 0u;
                u32 lc = 0u;
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    if(j == lj)
                    {
                        lpend = pending[j];
                        lc = cls[j];
                    }
                const u32 lcls = wave::lane(lc, hw::ffs(lpend) - 1u);
                RankNeed<M> need;
                u64 group[M];
                u32 K = 0u;
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    const bool in = ((pending[j] >> me) & 1u) != 0u && cls[j] == lcls;
                    group[j] = hw::ballot(in);
                    need.rank.rank[j] = in ? K + hw::rankBelow(group[j]) : kNoRank;
                    K += static_cast<u32>(hw::popc(group[j]));
                }
                if(K > cap)
                {
                    // the lowest `cap` ranks form the group; the rest stay pending
#pragma unroll
                    for(u32 j = 0; j < M; ++j)
                    {
                        if(need.rank.rank[j] >= cap)
                            need.rank.rank[j] = kNoRank;
                        group[j] = hw::ballot(need.rank.rank[j] != kNoRank);
                    }
                    K = cap;
                }
                need.K = K;
                need.got = 0u;
                u32 bMax = 0u, bMin = 0xffffffffu;
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    if(need.rank.rank[j] != kNoRank)
                    {
                        bMax = umax32(bMax, b[j]);
                        bMin = umin32(bMin, b[j]);
                    }
                const u32 bmax = wave::max(bMax);
                const u32 c = umax32(bmax, h.paddedMinChunk); // the largest member's chunk
                const u32 bmin = wave::min(bMin);
                const u32 maxcs = umin32(h.pagesize, umax32(h.wastefactor * bmin, h.paddedMinChunk));
                // chunks go straight to `result`: only served ranks are written
                allocGroup<M>(need, c, bmax, maxcs, ffb, result);
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    pending[j] &= ~group[j];
            }

            //! The leader slot of the next group: the lowest slot with a
            //! pending request, or M when none is left.
            template<u32 M>
            MALLOCMC_D static u32 leaderSlot(const u64 (&pending)[M])
            {
                u32 lj = M;
#pragma unroll
                for(u32 j = M; j-- > 0u;)
                    if(pending[j] != 0u)
                        lj = j;
                return lj;
            }

            //! Wave malloc of M requests per lane: slot j of lane l is the
            //! wave's request 64 j + l.  For callers that hand a wave more
            //! requests than it has lanes (the batch kernels): a size-class
            //! group then spans up to 64 M requests, served by the same scans
            //! and claims as malloc(), so the work a group pays once (its scan
            //! windows, claim round trips and wave reductions) is spread over
            //! more requests.  Groups are capped at aggregation x M requests.
            //! There is no distribution policy: XMallocSIMD coalesces the 64
            //! lanes of one call and keeps malloc().
            template<u32 M>
            MALLOCMC_D void mallocMany(const u32 (&bytesIn)[M], void* (&res)[M]) const
            {
                const u32 me = hw::laneid();
                // the access-block cursor, as in malloc()
                u32 ffb = 0u;
                if(me == 0u && h.accessblocks > 1u)
                    ffb = atom::load(&h.scalars->firstfreeblock);
                u32 b[M], cls[M];
                char* result[M];
                u64 pending[M];
                u32 failedPages = 0xffffffffu;
                beginSet<M>(bytesIn, b, result, pending, cls, failedPages);
                ffb = wave::lane(ffb, 0u);
                const u32 cap = t.aggregation * M;
                for(u32 lj = leaderSlot<M>(pending); lj != M; lj = leaderSlot<M>(pending))
                    serveGroup<M>(pending, cls, b, result, lj, cap, ffb);
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    res[j] = isOOM(result[j], b[j]) ? nullptr : result[j];
            }

            //! mallocMany over a stream of request sets in one group loop:
            //! `src.next(bytes)` fills the next set (false when there is none;
            //! the same answer on every lane) and `src.done(res)` takes the
            //! results of the set before.  A wave that serves several sets
            //! keeps the registers of one loop: a loop of mallocMany calls
            //! would keep every heap constant the call uses live across all of
            //! it.
            template<u32 M, class Src>
            MALLOCMC_D void mallocStream(Src& src) const
            {
                const u32 me = hw::laneid();
                u32 ffb = 0u;
                if(me == 0u && h.accessblocks > 1u)
                    ffb = atom::load(&h.scalars->firstfreeblock);
                ffb = wave::lane(ffb, 0u);
                const u32 cap = t.aggregation * M;
                u32 failedPages = 0xffffffffu;
                u32 b[M], cls[M];
                char* result[M];
                u64 pending[M];
                bool open = false; // a set is being served
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    b[j] = 0u;
                    cls[j] = 0xffffffffu;
                    result[j] = nullptr;
                    pending[j] = 0u;
                }
                for(;;)
                {
                    const u32 lj = leaderSlot<M>(pending);
                    if(lj != M)
                    {
                        serveGroup<M>(pending, cls, b, result, lj, cap, ffb);
                        continue;
                    }
                    // the set is served: hand over its results, take the next
                    if(open)
                    {
                        void* r[M];
#pragma unroll
                        for(u32 j = 0; j < M; ++j)
                            r[j] = isOOM(result[j], b[j]) ? nullptr : result[j];
                        src.done(r);
                    }
                    u32 bytesIn[M];
                    if(!src.next(bytesIn))
                        break;
                    open = true;
                    beginSet<M>(bytesIn, b, result, pending, cls, failedPages);
                }
            }

            //! Wave free: every lane passes its pointer or nullptr.  Lanes are
            //! grouped by page (and by 32-chunk segment on hierarchical pages)
            //! in registers first; then the leaders of all groups issue their
            //! atomics in the same vector instructions, so a wave pays a few
            //! round trips whatever the number of pages it touches.
            MALLOCMC_D void free(void* pv) const
            {
                const u32 me = hw::laneid();
                const u32 P = h.pagesize;
                char* p = static_cast<char*>(pv);
                u32 page = 0u, cs = 0u, inpage = 0u, q = 0u;
                bool simple = false;
                if(p != nullptr)
                {
                    page = h.pageOf(static_cast<u64>(p - h.pages));
                    cs = atom::load(h.pteChunksize(page));
                    inpage = static_cast<u32>(p - h.pageData(page));
                    const bool chunked = cs != 0u
                        && (cs < P || (cs == P && (h.flags & kFixPagesizeLeak)));
                    q = chunked ? h.pdiv(inpage, cs) : 0u;
                    simple = chunked && q * cs == inpage;
                    if(!simple)
                        h.destroy<WaveIds>(p); // multi-page, XMallocSIMD slice, stale pointer
                }
                const u32 idx = simple ? q : 0u;
                const bool hier = simple && cs <= h.hierarchyThreshold;
                const u32 seg = hier ? idx / 32u : 0u;

                // ---- grouping (registers only)
                bool pageLeader = false, segLeader = false;
                u32 pageK = 0u, pageBits = 0u, segBits = 0u;
                u64 pending = hw::ballot(simple);
                while(pending != 0u)
                {
                    const u32 leader = hw::ffs(pending) - 1u;
                    const u32 lpage = wave::lane(page, leader);
                    const bool member = simple && page == lpage;
                    const u64 group = hw::ballot(member);
                    // PTE bitmask bits to clear: chunk bits (flat) or segment hints (hierarchical)
                    const u32 bits = wave::bor(member ? (hier ? (1u << seg) : (1u << idx)) : 0u);
                    if(me == leader)
                    {
                        pageLeader = true;
                        pageK = hw::popc(group);
                        pageBits = bits;
                    }
                    if(wave::lane(static_cast<u32>(hier), leader) != 0u)
                    {
                        u64 segPending = group;
                        while(segPending != 0u)
                        {
                            const u32 l = hw::ffs(segPending) - 1u;
                            const u32 s = wave::lane(seg, l);
                            const bool inSeg = member && seg == s;
                            const u32 b = wave::bor(inSeg ? (1u << (idx % 32u)) : 0u);
                            if(me == l)
                            {
                                segLeader = true;
                                segBits = b;
                            }
                            segPending &= ~hw::ballot(inSeg);
                        }
                    }
                    pending &= ~group;
                }

                // ---- clear the chunk bits: every group at once
                if(segLeader)
                {
                    const u32 segmentsize = cs * 32u + 4u;
                    const u32 full = umin32(32u, h.pdiv(P, segmentsize));
                    const u32 additional = h.calcAdditional(full, segmentsize, cs);
                    u32* masks = h.onPageMasks(page, full + (additional > 0u ? 1u : 0u));
                    atom::and_(&masks[seg], ~segBits);
                }
                u32 old = 0u;
                if(pageLeader)
                {
                    atom::and_(h.pteBitmask(page), ~pageBits);
                    old = atom::sub(h.pteCount(page), pageK);
                    releaseFinish(page, releaseIssue(page, cs, old, pageK));
                }
            }

            //! One lane, after its group returned k chunks of `page` (chunk
            //! size cs) and the count read `old` before the decrement.
            //! Region reset when the count passes pagesize/2/c
            //! (Scatter.hpp:727-735), at least when the page empties: for
            //! c > pagesize/2 the reference's threshold is 0 and never fires,
            //! so failed scans ratchet every region to "full" and each later
            //! request scans the whole heap.  Hot words are tested first, so
            //! they are only written when they change.
            //! Then D4 for k chunks: the group that empties its page resets
            //! it.  The reference locks the count, clears the masks
            //! (Page::init), drops the chunk size, fences and unlocks
            //! (Scatter.hpp:686-716).  The side masks hold no bit of a freed
            //! chunk (each free clears its bits; a clear still in flight only
            //! keeps a chunk that no one owns out of use until it lands), and
            //! chunk size and count share one 64-bit word, so one CAS
            //! (cs, 0) -> (0, 0) resets the page if no one reserved it
            //! meanwhile.
            //! Two halves, so that a wave issues every group's loads and reset
            //! before it waits for any of them: `releaseIssue` reads the
            //! region's level and the access-block cursor (when the count
            //! crossed the threshold) and issues the page reset;
            //! `releaseFinish` writes the level and the cursor where they
            //! change.  (One after the other, the two tests cost a wave two
            //! dependent round trips more per free.)
            struct Release
            {
                u32 level;
                u32 cursor;
                bool crossed;
            };

            MALLOCMC_D Release releaseIssue(u32 page, u32 cs, u32 old, u32 k) const
            {
                const u32 thr = umax32(h.pdiv(h.pagesize / 2u, cs), 1u);
                Release r{0u, 0u, old >= thr && old - k < thr};
                if(r.crossed)
                {
                    r.level = atom::load(&h.regions[h.divR.div(page)]);
                    // one access block: the cursor is 0 and stays there
                    if(h.accessblocks > 1u)
                        r.cursor = atom::load(&h.scalars->firstfreeblock);
                }
                if(h.resetfreedpages && old == k)
                    atom::cas(h.ptePair(page), u64{cs}, u64{0});
                return r;
            }

            MALLOCMC_D void releaseFinish(u32 page, const Release& r) const
            {
                if(!r.crossed)
                    return;
                if(r.level != 0u)
                    atom::exch(&h.regions[h.divR.div(page)], 0u);
                const u32 block = h.divPpb.div(page);
                if(block < r.cursor)
                    atom::min(&h.scalars->firstfreeblock, block);
            }

            //! Wave free of M pointers per lane (slot j of lane l is the wave's
            //! request 64 j + l), grouped by page and segment over all 64 M
            //! slots, for callers that hand a wave more frees than it has
            //! lanes (the batch kernels): a page's frees meet in fewer groups,
            //! and a group's atomics are issued once for all its chunks.
            template<u32 M>
            MALLOCMC_D void freeMany(void* const (&pv)[M]) const
            {
                const u32 me = hw::laneid();
                const u32 P = h.pagesize;
                u32 page[M], cs[M], idx[M];
                bool simple[M];
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    page[j] = 0u;
                    cs[j] = 0u;
                    if(pv[j] != nullptr)
                    {
                        page[j] = h.pageOf(static_cast<u64>(static_cast<char*>(pv[j]) - h.pages));
                        cs[j] = atom::load(h.pteChunksize(page[j]));
                    }
                }
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    char* p = static_cast<char*>(pv[j]);
                    simple[j] = false;
                    idx[j] = 0u;
                    if(p != nullptr)
                    {
                        const u32 inpage = static_cast<u32>(p - h.pageData(page[j]));
                        const bool chunked = cs[j] != 0u
                            && (cs[j] < P || (cs[j] == P && (h.flags & kFixPagesizeLeak)));
                        const u32 q = chunked ? h.pdiv(inpage, cs[j]) : 0u;
                        simple[j] = chunked && q * cs[j] == inpage;
                        if(!simple[j])
                            h.destroy<WaveIds>(p); // multi-page, XMallocSIMD slice, stale pointer
                        idx[j] = simple[j] ? q : 0u;
                    }
                }

                // ---- grouping (registers only), slot by slot: a group is the
                // lanes of one slot j that free chunks of one page, led by its
                // lowest lane.  (Groups across the slots, one per page, took
                // more instructions per page than the extra group costs where
                // a page's frees span two slots; the callers hand a wave its
                // frees sorted by page index mod 256, so a page's frees lie in
                // one slot or two neighbouring ones.)
                bool pageLeader[M], segLeader[M];
                u32 pageK[M], pageBits[M], segBits[M];
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    pageLeader[j] = false;
                    segLeader[j] = false;
                    pageK[j] = 0u;
                    pageBits[j] = 0u;
                    segBits[j] = 0u;
                    u64 pending = hw::ballot(simple[j]);
                    while(pending != 0u)
                    {
                        const u32 leader = hw::ffs(pending) - 1u;
                        const u32 lpage = wave::lane(page[j], leader);
                        // every member shares the page, hence its chunk size
                        const bool hier = wave::lane(cs[j], leader) <= h.hierarchyThreshold;
                        const bool member = ((pending >> me) & 1u) != 0u && page[j] == lpage;
                        const u64 group = hw::ballot(member);
                        // PTE bitmask bits to clear: chunk bits (flat) or segment hints (hierarchical)
                        const u32 bits
                            = wave::bor(member ? (hier ? (1u << (idx[j] / 32u)) : (1u << idx[j])) : 0u);
                        if(me == leader)
                        {
                            pageLeader[j] = true;
                            pageK[j] = static_cast<u32>(hw::popc(group));
                            pageBits[j] = bits;
                        }
                        if(hier)
                        {
                            for(u64 segPending = group; segPending != 0u;)
                            {
                                const u32 sl = hw::ffs(segPending) - 1u;
                                const u32 seg = wave::lane(idx[j], sl) / 32u;
                                const bool inSeg = ((segPending >> me) & 1u) != 0u && idx[j] / 32u == seg;
                                const u32 b = wave::bor(inSeg ? 1u << (idx[j] % 32u) : 0u);
                                if(me == sl)
                                {
                                    segLeader[j] = true;
                                    segBits[j] = b;
                                }
                                segPending &= ~hw::ballot(inSeg);
                            }
                        }
                        pending &= ~group;
                    }
                }

                // ---- clear the chunk bits and return the counts: every group at once
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    if(segLeader[j])
                    {
                        const u32 segmentsize = cs[j] * 32u + 4u;
                        const u32 full = umin32(32u, h.pdiv(P, segmentsize));
                        const u32 additional = h.calcAdditional(full, segmentsize, cs[j]);
                        u32* masks = h.onPageMasks(page[j], full + (additional > 0u ? 1u : 0u));
                        atom::and_(&masks[idx[j] / 32u], ~segBits[j]);
                    }
                u32 old[M];
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    old[j] = 0u;
                    if(pageLeader[j])
                    {
                        atom::and_(h.pteBitmask(page[j]), ~pageBits[j]);
                        old[j] = atom::sub(h.pteCount(page[j]), pageK[j]);
                    }
                }
                Release rel[M];
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                {
                    rel[j] = Release{0u, 0u, false};
                    if(pageLeader[j])
                        rel[j] = releaseIssue(page[j], cs[j], old[j], pageK[j]);
                }
#pragma unroll
                for(u32 j = 0; j < M; ++j)
                    releaseFinish(page[j], rel[j]);
            }
        };
    } // namespace detail
} // namespace mallocMC
