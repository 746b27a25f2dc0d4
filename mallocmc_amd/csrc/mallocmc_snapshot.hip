// Heap snapshots and their host-side verification.
// Declarations and format: include/mallocmc_snapshot.h.
#include "mallocmc_snapshot.h"

#include "mmc_internal.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace mallocMC;
using namespace mallocMC::detail;

namespace
{
    constexpr uint32_t kSnapshotVersion = 1u;

    uint64_t align8(uint64_t x)
    {
        return (x + 7u) & ~uint64_t{7};
    }

    struct Sections
    {
        uint64_t ptes, regions, masks, payload, total;
    };

    Sections sectionsFor(const mmc_geometry& g, uint32_t pagesize, uint32_t flags)
    {
        Sections s{};
        s.ptes = align8(sizeof(mmc_snapshot_header));
        s.regions = align8(s.ptes + 12u * static_cast<uint64_t>(g.numpages));
        s.masks = align8(s.regions + 4u * static_cast<uint64_t>(g.numregions));
        uint64_t end = align8(s.masks + 4u * static_cast<uint64_t>(g.sideMaskWords) * g.numpages);
        s.payload = 0u;
        if(flags & MMC_SNAPSHOT_PAYLOAD)
        {
            s.payload = end;
            end = align8(end + static_cast<uint64_t>(g.numpages) * pagesize);
        }
        s.total = end;
        return s;
    }

    uint32_t popc32(uint32_t x)
    {
        return static_cast<uint32_t>(__builtin_popcount(x));
    }

    //! Page layout of a chunk size, the host restatement of the capacity rules
    //! of Scatter.hpp:365-383, 494-510 (the same arithmetic as ScatterHeap).
    struct HostLayout
    {
        bool hier;
        uint32_t full, additional, segments, capacity;
    };

    HostLayout hostLayout(uint32_t P, uint32_t ht, uint32_t cs)
    {
        HostLayout L{};
        L.hier = cs <= ht;
        if(L.hier)
        {
            const uint32_t seg = cs * 32u + 4u;
            L.full = std::min(32u, P / seg);
            if(L.full < 32u)
            {
                const uint64_t used = static_cast<uint64_t>(L.full) * seg + 4u;
                L.additional = used >= P ? 0u : std::min(31u, static_cast<uint32_t>((P - used) / cs));
            }
            L.segments = L.full + (L.additional > 0u ? 1u : 0u);
            L.capacity = L.full * 32u + L.additional;
        }
        else
        {
            L.segments = 0u;
            L.capacity = std::min(P / cs, 32u);
        }
        return L;
    }

    uint32_t spotMask(uint32_t spots)
    {
        return spots >= 32u ? ~0u : ((1u << spots) - 1u);
    }

    struct Checker
    {
        mmc_verify_report* r;

        void error(uint32_t code, uint64_t where, const char* fmt, ...)
        {
            if(r->errors++ == 0u)
            {
                r->first_error_code = code;
                r->first_error_page = where;
                va_list ap;
                va_start(ap, fmt);
                std::vsnprintf(r->message, sizeof(r->message), fmt, ap);
                va_end(ap);
            }
        }
    };
} // namespace

extern "C" {

int mmc_snapshot_bytes(const mmc_heap* heap, uint32_t flags, uint64_t* bytes)
{
    if(!heap || !bytes)
        return mmc::fail(MMC_EINVAL, "null argument");
    mmc_geometry g{};
    if(int rc = mmc_get_geometry(heap, &g))
        return rc;
    *bytes = sectionsFor(g, heap->cfg.pagesize, flags).total;
    return MMC_OK;
}

int mmc_snapshot(mmc_heap* heap, void* dst, uint64_t capacity, uint32_t flags, void* stream)
{
    if(!heap || !dst)
        return mmc::fail(MMC_EINVAL, "null argument");
    mmc_snapshot_header hdr{};
    std::memcpy(hdr.magic, MMC_SNAPSHOT_MAGIC, 8);
    hdr.version = kSnapshotVersion;
    hdr.flags = flags & MMC_SNAPSHOT_PAYLOAD;
    hdr.cfg = heap->cfg;
    if(int rc = mmc_get_geometry(heap, &hdr.geom))
        return rc;
    hdr.memsize = heap->memsize;
    const Sections s = sectionsFor(hdr.geom, heap->cfg.pagesize, hdr.flags);
    if(capacity < s.total)
        return mmc::fail(MMC_EINVAL, "snapshot buffer too small (see mmc_snapshot_bytes)");
    hdr.off_ptes = s.ptes;
    hdr.off_regions = s.regions;
    hdr.off_masks = s.masks;
    hdr.off_payload = s.payload;
    hdr.total_bytes = s.total;
    char* out = static_cast<char*>(dst);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const ScatterHeap& v = heap->view;
    ScatterScalars sc{};
    hipError_t err = hipMemcpyAsync(&sc, heap->scalars, sizeof(sc), hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipMemcpyAsync(out + s.ptes, v.ptes, 12u * static_cast<uint64_t>(hdr.geom.numpages),
                             hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipMemcpyAsync(out + s.regions, v.regions, 4u * static_cast<uint64_t>(hdr.geom.numregions),
                             hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipMemcpyAsync(out + s.masks, v.masks,
                             4u * static_cast<uint64_t>(hdr.geom.sideMaskWords) * hdr.geom.numpages,
                             hipMemcpyDeviceToHost, st);
    if(err == hipSuccess && (hdr.flags & MMC_SNAPSHOT_PAYLOAD))
        err = hipMemcpyAsync(out + s.payload, v.pages,
                             static_cast<uint64_t>(hdr.geom.numpages) * heap->cfg.pagesize,
                             hipMemcpyDeviceToHost, st);
    if(err == hipSuccess)
        err = hipStreamSynchronize(st);
    if(err != hipSuccess)
        return mmc::hipFail(err, "snapshot copy");
    // odd pages keep [bitmask, chunksize, count] on the device (chunk size
    // and count share an 8-byte word); the snapshot holds every PTE as
    // [chunksize, count, bitmask]
    for(uint64_t p = 1; p < hdr.geom.numpages; p += 2)
    {
        uint32_t w[3];
        char* e = out + s.ptes + 12u * p;
        std::memcpy(w, e, 12);
        const uint32_t canon[3] = {w[1], w[2], w[0]};
        std::memcpy(e, canon, 12);
    }
    hdr.firstFreePageBased = sc.firstFreePageBased;
    hdr.firstfreeblock = sc.firstfreeblock;
    hdr.pagebasedMutex = sc.pagebasedMutex;
    std::memcpy(out, &hdr, sizeof(hdr));
    return MMC_OK;
}

int mmc_snapshot_verify(const void* snapshot, uint64_t bytes, const uint64_t* live_offsets,
                        const uint32_t* live_sizes, uint64_t n_live, uint32_t flags, mmc_verify_report* report)
{
    if(!snapshot || !report)
        return mmc::fail(MMC_EINVAL, "null argument");
    std::memset(report, 0, sizeof(*report));
    Checker ck{report};
    if(bytes < sizeof(mmc_snapshot_header))
    {
        ck.error(10, 0, "snapshot shorter than its header");
        return MMC_OK;
    }
    mmc_snapshot_header hdr;
    std::memcpy(&hdr, snapshot, sizeof(hdr));
    const mmc_geometry& g = hdr.geom;
    const uint32_t P = hdr.cfg.pagesize;
    if(std::memcmp(hdr.magic, MMC_SNAPSHOT_MAGIC, 8) != 0 || hdr.version != kSnapshotVersion || P == 0u
       || hdr.total_bytes > bytes || g.sideMaskWords > 32u)
    {
        ck.error(10, 0, "not a version-1 heap snapshot, or truncated");
        return MMC_OK;
    }
    const Sections s = sectionsFor(g, P, hdr.flags);
    if(s.total != hdr.total_bytes || s.ptes != hdr.off_ptes || s.masks != hdr.off_masks)
    {
        ck.error(10, 0, "section offsets do not match the geometry");
        return MMC_OK;
    }
    const char* base = static_cast<const char*>(snapshot);
    auto pte = [&](uint32_t p, uint32_t f) {
        uint32_t v;
        std::memcpy(&v, base + s.ptes + 12u * static_cast<uint64_t>(p) + 4u * f, 4);
        return v;
    };
    auto mask = [&](uint32_t p, uint32_t w) {
        uint32_t v;
        std::memcpy(&v, base + s.masks + 4u * (static_cast<uint64_t>(p) * g.sideMaskWords + w), 4);
        return v;
    };
    const uint32_t ht = g.hierarchyThreshold;
    const bool fixLeak = (hdr.cfg.flags & 1u) != 0u;
    (void) fixLeak;

    // ---- scalars
    if(hdr.firstFreePageBased >= g.numpages || (g.accessblocks && hdr.firstfreeblock >= g.accessblocks)
       || hdr.pagebasedMutex > 1u)
        ck.error(5, 0, "heap scalars out of range: firstFreePageBased %u firstfreeblock %u mutex %u",
                 hdr.firstFreePageBased, hdr.firstfreeblock, hdr.pagebasedMutex);

    // ---- pages
    std::vector<uint32_t> runStart(g.numpages, 0xffffffffu); // page -> first page of its run
    report->pages = g.numpages;
    for(uint32_t p = 0; p < g.numpages; ++p)
    {
        const uint32_t cs = pte(p, 0), cnt = pte(p, 1), bits = pte(p, 2);
        uint32_t anyMask = 0u;
        for(uint32_t w = 0; w < g.sideMaskWords; ++w)
            anyMask |= mask(p, w);
        if(cs == 0u)
        {
            ++report->free_pages;
            if(cnt != 0u || bits != 0u || anyMask != 0u)
                ck.error(1, p, "page %u is free but has count %u, bitmask %08x, masks %08x", p, cnt, bits, anyMask);
            continue;
        }
        if(cs > P)
        {
            // first page of a run, or a continuation of the run before it
            if(p > 0u && runStart[p - 1u] != 0xffffffffu && pte(runStart[p - 1u], 0) == cs
               && static_cast<uint64_t>(p - runStart[p - 1u]) * P < cs)
                runStart[p] = runStart[p - 1u];
            else
            {
                runStart[p] = p;
                ++report->runs;
                const uint64_t k = (static_cast<uint64_t>(cs) + P - 1u) / P;
                if(p + k > g.numpages)
                    ck.error(4, p, "run of %u bytes at page %u passes the last page", cs, p);
                else
                    for(uint64_t q = p + 1u; q < p + k; ++q)
                        if(pte(static_cast<uint32_t>(q), 0) != cs)
                        {
                            ck.error(4, p, "run of %u bytes at page %u: page %llu has chunk size %u", cs, p,
                                     static_cast<unsigned long long>(q), pte(static_cast<uint32_t>(q), 0));
                            break;
                        }
            }
            ++report->run_pages;
            if(cnt != 0u || bits != 0u || anyMask != 0u)
                ck.error(1, p, "page %u of a run has count %u, bitmask %08x, masks %08x", p, cnt, bits, anyMask);
            continue;
        }
        ++report->chunked_pages;
        const HostLayout L = hostLayout(P, ht, cs);
        uint32_t used = 0u;
        if(L.hier)
        {
            for(uint32_t w = 0; w < g.sideMaskWords; ++w)
            {
                const uint32_t spots = w < L.full ? 32u : (w < L.segments ? L.additional : 0u);
                const uint32_t m = mask(p, w);
                if(m & ~spotMask(spots))
                    ck.error(2, p, "page %u (chunk %u): mask word %u has bits %08x beyond its %u chunks", p, cs, w, m,
                             spots);
                used += popc32(m & spotMask(spots));
                const bool hint = w < 32u && ((bits >> w) & 1u) != 0u;
                if(hint && w < L.segments && (m & spotMask(spots)) != spotMask(spots))
                    ++report->stale_hints;
            }
            if(L.segments < 32u && (bits >> L.segments) != 0u)
                ck.error(2, p, "page %u (chunk %u): segment hints %08x beyond its %u segments", p, cs, bits,
                         L.segments);
        }
        else
        {
            if(anyMask != 0u)
                ck.error(2, p, "flat page %u (chunk %u) has segment masks %08x", p, cs, anyMask);
            if(bits & ~spotMask(L.capacity))
                ck.error(2, p, "page %u (chunk %u): bitmask %08x beyond its %u chunks", p, cs, bits, L.capacity);
            used = popc32(bits & spotMask(L.capacity));
        }
        report->used_chunks += used;
        if(cnt != used)
            ck.error(3, p, "page %u (chunk %u): count %u but %u chunks marked in use", p, cs, cnt, used);
        if(cnt == 0u && hdr.cfg.resetfreedpages)
            ++report->empty_formatted;
    }

    // ---- live allocations
    if(n_live == 0u || !live_offsets)
        return MMC_OK;
    const uint64_t pageBytes = static_cast<uint64_t>(g.numpages) * P;
    std::vector<uint64_t> keys; // chunk key (page * 1024 + index) or run key (2^63 | page)
    keys.reserve(n_live);
    for(uint64_t i = 0; i < n_live; ++i)
    {
        const uint64_t off = live_offsets[i];
        const uint32_t sz = live_sizes ? live_sizes[i] : 0u;
        ++report->live_checked;
        if(off >= pageBytes)
        {
            ck.error(6, i, "live %llu: offset %llu outside the page area", static_cast<unsigned long long>(i),
                     static_cast<unsigned long long>(off));
            continue;
        }
        const uint32_t p = static_cast<uint32_t>(off / P);
        const uint32_t inpage = static_cast<uint32_t>(off % P);
        const uint32_t cs = pte(p, 0);
        if(cs == 0u)
        {
            ck.error(6, i, "live %llu: offset %llu on free page %u", static_cast<unsigned long long>(i),
                     static_cast<unsigned long long>(off), p);
            continue;
        }
        if(cs > P)
        {
            if(runStart[p] != p || inpage != 0u)
            {
                ck.error(6, i, "live %llu: offset %llu is not the start of a run", static_cast<unsigned long long>(i),
                         static_cast<unsigned long long>(off));
                continue;
            }
            if(sz > cs)
                ck.error(7, i, "live %llu: %u bytes in a run of %u", static_cast<unsigned long long>(i), sz, cs);
            keys.push_back((uint64_t{1} << 63) | p);
            continue;
        }
        const HostLayout L = hostLayout(P, ht, cs);
        const uint32_t idx = inpage / cs;
        const uint32_t within = inpage % cs;
        if((within != 0u && !(flags & MMC_VERIFY_LIVE_INTERIOR)) || idx >= L.capacity)
        {
            ck.error(6, i, "live %llu: offset %llu is not a chunk of page %u (chunk %u)",
                     static_cast<unsigned long long>(i), static_cast<unsigned long long>(off), p, cs);
            continue;
        }
        const bool set = L.hier ? ((mask(p, idx / 32u) >> (idx % 32u)) & 1u) != 0u : ((pte(p, 2) >> idx) & 1u) != 0u;
        if(!set)
            ck.error(7, i, "live %llu: chunk %u of page %u is not marked in use", static_cast<unsigned long long>(i),
                     idx, p);
        if(sz > cs - within)
            ck.error(7, i, "live %llu: %u bytes at %u into a chunk of %u", static_cast<unsigned long long>(i), sz,
                     within, cs);
        keys.push_back(static_cast<uint64_t>(p) * 1024u + idx);
    }
    std::sort(keys.begin(), keys.end());
    if(!(flags & MMC_VERIFY_LIVE_INTERIOR))
        for(size_t i = 1; i < keys.size(); ++i)
            if(keys[i] == keys[i - 1u])
            {
                const bool run = (keys[i] >> 63) != 0u;
                const uint64_t k = keys[i] & ~(uint64_t{1} << 63);
                ck.error(8, run ? k : k / 1024u, run ? "two live allocations on the run at page %llu"
                                                     : "two live allocations on chunk %llu of page %llu",
                         run ? static_cast<unsigned long long>(k) : static_cast<unsigned long long>(k % 1024u),
                         static_cast<unsigned long long>(k / 1024u));
                break;
            }
    if(flags & MMC_VERIFY_LIVE_COMPLETE)
    {
        // every page's count equals its live chunks, every run has its allocation
        std::vector<uint32_t> liveOnPage(g.numpages, 0u);
        std::vector<uint8_t> runLive(g.numpages, 0u);
        for(size_t i = 0; i < keys.size(); ++i)
        {
            if(i > 0u && keys[i] == keys[i - 1u])
                continue;
            if(keys[i] >> 63)
                runLive[keys[i] & 0xffffffffu] = 1u;
            else
                ++liveOnPage[keys[i] / 1024u];
        }
        for(uint32_t p = 0; p < g.numpages; ++p)
        {
            const uint32_t cs = pte(p, 0);
            if(cs != 0u && cs <= P && pte(p, 1) != liveOnPage[p])
            {
                ck.error(9, p, "page %u: count %u but %u live allocations", p, pte(p, 1), liveOnPage[p]);
                break;
            }
            if(cs > P && runStart[p] == p && !runLive[p])
            {
                ck.error(9, p, "run at page %u has no live allocation", p);
                break;
            }
        }
    }
    return MMC_OK;
}

} // extern "C"
