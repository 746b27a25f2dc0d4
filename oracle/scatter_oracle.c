/*
 * CPU oracle for the ScatterAlloc device heap -- TEST INFRASTRUCTURE ONLY.
 * See scatter_oracle.h for scope, execution model and divergences D1-D3.
 *
 * Memory model: the pool is a real byte array laid out exactly as the
 * reference lays it out (Scatter.hpp:999-1014): numpages pages, then one
 * 12-byte PTE {chunksize, count, bitmask} per page, then one uint32 fill
 * level per region.  On-page bitmasks of hierarchical pages live in the last
 * 4*segments bytes of the page (Scatter.hpp:315-318).
 */
#define _GNU_SOURCE
#include "scatter_oracle.h"

#include <omp.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <time.h>

typedef uint32_t u32;
typedef uint64_t u64;

/* Test-only mutants.  The oracle's pinning tests (tests/test_oracle_kats.py)
 * build this file once more per ORC_MUTANT value; each value breaks one
 * stateful heuristic of the walk, the way a plausible misreading of
 * Scatter.hpp would, so the tests can show that their known-answer streams
 * catch it.  0, the default build, is the faithful restatement, and nothing
 * but those tests loads a mutant build. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define ORC_MUT_NO_REGION_INCREMENT 1  /* drop the fill-level CAS, :612-617 */
#define ORC_MUT_INCREMENT_BELOW_R 2    /* cap it at fill + 1 < R instead of <= R */
#define ORC_MUT_NO_REGION_RESET 3      /* drop the exchange to 0, :727-729 */
#define ORC_MUT_RESET_WHEN_EMPTY 4     /* reset when the page empties, not at P/2/c */
#define ORC_MUT_NO_BLOCK_ADVANCE 5     /* drop the cursor advance, :620-629 */
#define ORC_MUT_ADVANCE_ALWAYS 6       /* advance without the `>` guard */
#define ORC_MUT_NO_BLOCK_MIN 7         /* drop the atomicMin of the cursor, :733-734 */
#define ORC_MUT_SINGLE_PASS 8          /* no second pass, :634 */
#define ORC_MUT_CHECKLEVEL_R 9         /* first pass gated at R instead of 3R/4, :553 */
#define ORC_MUT_NO_FFPB_CAS 10         /* drop the _firstFreePageBased CAS, :796-801 */
#define ORC_MUT_NO_FFPB_MAX 11         /* drop its atomicMax on free, :885 */
#define MUT(name) (ORC_MUTANT == ORC_MUT_##name)

struct orc_heap {
    orc_config cfg;
    /* constants, Scatter.hpp:173-187 */
    u32 HT;
    u32 minChunkSize;
    u32 maxOnPageMasks;
    /* state, Scatter.hpp:263-271 */
    uint8_t* pool;
    u64 memsize;
    u32* ptes; /* 3 words per page */
    u32* regions;
    u32 numpages;
    u32 numregions;
    u32 accessblocks;
    u32 pagebasedMutex;
    u32 firstFreePageBased;
    u32 firstfreeblock;
    u32 clamped;
    int atomic_mode;
};

/* ---------------------------------------------------------------- atomics */
/* alpaka's CPU grid atomics are AtomicStdLibLock<16>
 * (alpaka/include/alpaka/atomic/AtomicCpu.hpp:9-21): a std::mutex chosen by
 * (address / sizeof(T)) & 15 (AtomicStdLibLock.hpp:41-72). */
static pthread_mutex_t g_locks[16] = {
    PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER,
    PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER,
    PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER,
    PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER, PTHREAD_MUTEX_INITIALIZER};

static inline pthread_mutex_t* lock_for(const void* p, size_t sz)
{
    return &g_locks[((uintptr_t) p / sz) & 15u];
}

enum { OP_ADD, OP_SUB, OP_OR, OP_AND, OP_EXCH, OP_MIN, OP_MAX };

static inline u32 rmw_plain(u32* p, int op, u32 v)
{
    u32 old = *p;
    switch(op)
    {
    case OP_ADD: *p = old + v; break;
    case OP_SUB: *p = old - v; break;
    case OP_OR: *p = old | v; break;
    case OP_AND: *p = old & v; break;
    case OP_EXCH: *p = v; break;
    case OP_MIN: *p = old < v ? old : v; break;
    case OP_MAX: *p = old > v ? old : v; break;
    }
    return old;
}

static inline u32 rmw(orc_heap* h, u32* p, int op, u32 v)
{
    if(h->atomic_mode == ORC_ATOMIC_MUTEX16)
    {
        pthread_mutex_t* m = lock_for(p, sizeof(u32));
        pthread_mutex_lock(m);
        u32 old = rmw_plain(p, op, v);
        pthread_mutex_unlock(m);
        return old;
    }
    switch(op)
    {
    case OP_ADD: return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
    case OP_SUB: return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST);
    case OP_OR: return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
    case OP_AND: return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST);
    case OP_EXCH: return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST);
    case OP_MIN:
    {
        u32 old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
        while(old > v && !__atomic_compare_exchange_n(p, &old, v, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST))
        {
        }
        return old;
    }
    default:
    {
        u32 old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
        while(old < v && !__atomic_compare_exchange_n(p, &old, v, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST))
        {
        }
        return old;
    }
    }
}

static inline u32 cas(orc_heap* h, u32* p, u32 cmp, u32 val)
{
    if(h->atomic_mode == ORC_ATOMIC_MUTEX16)
    {
        pthread_mutex_t* m = lock_for(p, sizeof(u32));
        pthread_mutex_lock(m);
        u32 old = *p;
        if(old == cmp)
            *p = val;
        pthread_mutex_unlock(m);
        return old;
    }
    u32 expected = cmp;
    __atomic_compare_exchange_n(p, &expected, val, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    return expected;
}

/* volatile-style plain read of shared state (the reference reads PTE fields
 * through `volatile PTE*`, Scatter.hpp:263-264) */
static inline u32 vload(const u32* p)
{
    return __atomic_load_n(p, __ATOMIC_RELAXED);
}

static inline void vstore(u32* p, u32 v)
{
    __atomic_store_n(p, v, __ATOMIC_RELAXED);
}

static inline void fence_device(void)
{
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
}

/* --------------------------------------------------------- bit utilities */
/* CPU fallbacks of mallocMC_utils.hpp:205-245: ffs is 1-based, 0 for 0. */
static inline u32 ffs32(u32 m)
{
    if(m == 0)
        return 0;
    return (u32) __builtin_ctz(m) + 1u;
}

static inline u32 popc32(u32 m)
{
    return (u32) __builtin_popcount(m);
}

static inline u32 divup32(u32 a, u32 b)
{
    return (a + b - 1u) / b;
}

static inline u32 umin(u32 a, u32 b)
{
    return a < b ? a : b;
}

static inline u32 umax(u32 a, u32 b)
{
    return a > b ? a : b;
}

/* Shrink::applyPadding (Shrink.hpp:132-137); dataAlignment 1 == Noop. */
static inline u32 apply_padding(const orc_heap* h, u32 bytes)
{
    const u32 clear = h->cfg.dataAlignment - 1u;
    return (bytes + clear) & ~clear;
}

#define PTE_CHUNKSIZE(h, p) (&(h)->ptes[3u * (u64) (p) + 0u])
#define PTE_COUNT(h, p) (&(h)->ptes[3u * (u64) (p) + 1u])
#define PTE_BITMASK(h, p) (&(h)->ptes[3u * (u64) (p) + 2u])

static inline uint8_t* page_data(const orc_heap* h, u32 page)
{
    return h->pool + (u64) page * h->cfg.pagesize;
}

/* Page::init (Scatter.hpp:249-258): zero the last 4*maxOnPageMasks bytes. */
static void page_init(const orc_heap* h, u32 page)
{
    u32* w = (u32*) (page_data(h, page) + h->cfg.pagesize - 4u * h->maxOnPageMasks);
    for(u32 i = 0; i < h->maxOnPageMasks; ++i)
        vstore(&w[i], 0u);
}

/* onPageMasksPosition (Scatter.hpp:315-318) */
static inline u32* on_page_masks(const orc_heap* h, u32 page, u32 nMasks)
{
    return (u32*) (page_data(h, page) + h->cfg.pagesize - 4u * nMasks);
}

/* nextspot (Scatter.hpp:293-304): next zero bit to the left of spot, wrapping
 * within `spots` bits. */
static inline u32 nextspot(u32 bitfield, u32 spot, u32 spots)
{
    const u32 low = (spot + 1u) == 32u ? 0u : (bitfield >> (spot + 1u));
    const u32 high = bitfield << (spots - (spot + 1u));
    const u32 sel = spots == 32u ? ~0u : ((1u << spots) - 1u);
    const u32 rot = (high | low) & sel;
    const u32 step = ffs32(~rot);
    return (spot + step) % spots;
}

/* usespot (Scatter.hpp:329-347), randInit() = laneid() = 0 on the CPU. */
static int usespot(orc_heap* h, u32* bitfield, u32 spots)
{
    u32 spot = 0u % spots;
    for(;;)
    {
        const u32 mask = 1u << spot;
        const u32 old = rmw(h, bitfield, OP_OR, mask);
        if((old & mask) == 0u)
            return (int) spot;
        if(popc32(old) >= spots)
            return -1;
        spot = nextspot(old, spot, spots);
    }
}

/* calcAdditionalChunks (Scatter.hpp:365-383), evaluated in signed arithmetic
 * (the remainder P - full*(32c+4) is >= 4 for every 16-aligned chunk size at
 * pagesize 4K/16K/64K, so signed and unsigned readings agree there). */
static inline u32 calc_additional(const orc_heap* h, u32 fullsegments, u32 segmentsize, u32 chunksize)
{
    if(fullsegments == 32u)
        return 0u;
    long long rem = (long long) h->cfg.pagesize - (long long) fullsegments * (long long) segmentsize - 4ll;
    if(rem < 0)
        rem = 0;
    const long long add = rem / (long long) chunksize;
    return (u32) (add < 31 ? add : 31);
}

/* addChunkHierarchy (Scatter.hpp:397-421), laneid() = 0. */
static uint8_t* add_chunk_hierarchy(orc_heap* h, u32 chunksize, u32 fullsegments, u32 additional, u32 page)
{
    const u32 segments = fullsegments + (additional > 0u ? 1u : 0u);
    u32 spot = 0u % segments;
    const u32 mask = vload(PTE_BITMASK(h, page));
    if((mask & (1u << spot)) != 0u)
        spot = nextspot(mask, spot, segments);
    const u32 tries = segments - popc32(mask);
    u32* masks = on_page_masks(h, page, segments);
    for(u32 i = 0; i < tries; ++i)
    {
        const int hspot = usespot(h, &masks[spot], spot < fullsegments ? 32u : additional);
        if(hspot != -1)
            return page_data(h, page) + (u64) (32u * spot + (u32) hspot) * chunksize;
        rmw(h, PTE_BITMASK(h, page), OP_OR, 1u << spot);
        spot = nextspot(mask, spot, segments);
    }
    return NULL;
}

/* addChunkNoHierarchy (Scatter.hpp:432-443) */
static uint8_t* add_chunk_no_hierarchy(orc_heap* h, u32 chunksize, u32 page, u32 spots)
{
    const int spot = usespot(h, PTE_BITMASK(h, page), spots);
    if(spot == -1)
        return NULL;
    return page_data(h, page) + (u64) (u32) spot * chunksize;
}

/* tryUsePage (Scatter.hpp:454-518); the range check is the lambda of
 * allocChunked (:589-591). */
static uint8_t* try_use_page(orc_heap* h, u32 page, u32 chunksize, u32 bytes, u32 maxchunksize)
{
    uint8_t* chunk = NULL;
    const u32 P = h->cfg.pagesize;
    const u32 filllevel = rmw(h, PTE_COUNT(h, page), OP_ADD, 1u);
    int tryAlloc = !h->cfg.resetfreedpages;
    if(h->cfg.resetfreedpages && filllevel < P)
    {
        const u32 oldcs = cas(h, PTE_CHUNKSIZE(h, page), 0u, chunksize);
        if(oldcs == 0u || (oldcs >= bytes && oldcs <= maxchunksize))
            tryAlloc = 1;
        if(oldcs != 0u)
            chunksize = oldcs;
    }
    if(tryAlloc)
    {
        if(chunksize <= h->HT)
        {
            const u32 segmentsize = chunksize * 32u + 4u;
            const u32 fullsegments = umin(32u, P / segmentsize);
            const u32 additional = calc_additional(h, fullsegments, segmentsize, chunksize);
            if(filllevel < fullsegments * 32u + additional)
                chunk = add_chunk_hierarchy(h, chunksize, fullsegments, additional, page);
        }
        else
        {
            const u32 chunksinpage = umin(P / chunksize, 32u);
            if(filllevel < chunksinpage)
                chunk = add_chunk_no_hierarchy(h, chunksize, page, chunksinpage);
        }
    }
    if(chunk == NULL)
        rmw(h, PTE_COUNT(h, page), OP_SUB, 1u);
    return chunk;
}

/* allocChunked (Scatter.hpp:527-637) with smid() = warpid() = 0, warpSize = 1. */
static uint8_t* alloc_chunked(orc_heap* h, u32 bytes)
{
    const u32 P = h->cfg.pagesize;
    const u32 R = h->cfg.regionsize;
    const u32 paddedMin = apply_padding(h, h->minChunkSize);
    const u32 minAllocation = umax(bytes, paddedMin);
    const u32 numpages = h->numpages;
    const u32 pagesperblock = numpages / h->accessblocks;
    const u32 reloff = 1u * minAllocation / P;
    const u32 smid = 0u, warpid = 0u;
    const u32 start_in_block
        = (minAllocation * h->cfg.hashingK + h->cfg.hashingDistMP * smid
           + (h->cfg.hashingDistWP + h->cfg.hashingDistWPRel * reloff) * warpid)
        % pagesperblock;
    const u32 maxchunksize = umin(P, umax(h->cfg.wastefactor * bytes, paddedMin));
    const u32 global_start = start_in_block + vload(&h->firstfreeblock) * pagesperblock;

    u32 checklevel = MUT(CHECKLEVEL_R) ? R : R * 3u / 4u;
    for(u32 finder = 0; finder < (MUT(SINGLE_PASS) ? 1u : 2u); ++finder)
    {
        u32 global_page = global_start;
        do
        {
            const u32 region = global_page / R;
            const u32 fill = vload(&h->regions[region]);
            const u32 region_offset = region * R;
            if(fill < checklevel)
            {
                u32 pir = global_page;
                do
                {
                    const u32 before = cas(h, PTE_CHUNKSIZE(h, pir), 0u, minAllocation);
                    u32 use = 0u;
                    if(before == 0u)
                        use = minAllocation;
                    else if(before >= bytes && before <= maxchunksize)
                        use = before;
                    if(use != 0u)
                    {
                        uint8_t* res = try_use_page(h, pir, use, bytes, maxchunksize);
                        if(res != NULL)
                            return res;
                    }
                    pir = region_offset + ((pir + 1u) % R);
                } while(pir != global_page);
                if(!MUT(NO_REGION_INCREMENT) && fill + 1u <= R - (MUT(INCREMENT_BELOW_R) ? 1u : 0u))
                    cas(h, &h->regions[region], fill, fill + 1u);
            }
            global_page = (global_page + R) % numpages;
            if(!MUT(NO_BLOCK_ADVANCE) && global_page % pagesperblock == 0u)
            {
                const u32 abid = global_page / pagesperblock;
                if(MUT(ADVANCE_ALWAYS) || abid > vload(&h->firstfreeblock))
                    vstore(&h->firstfreeblock, abid);
            }
        } while(global_page != global_start);
        checklevel = R + 1u;
    }
    return NULL;
}

/* deallocChunked (Scatter.hpp:646-736) with warpid() + laneid() == 0. */
static void dealloc_chunked(orc_heap* h, uint8_t* mem, u32 page, u32 chunksize)
{
    const u32 P = h->cfg.pagesize;
    const u32 inpage = (u32) (mem - page_data(h, page));
    if(chunksize <= h->HT)
    {
        const u32 segmentsize = chunksize * 32u + 4u;
        const u32 fullsegments = umin(32u, P / segmentsize);
        const u32 additional = calc_additional(h, fullsegments, segmentsize, chunksize);
        const u32 segment = inpage / (chunksize * 32u);
        const u32 within = (inpage - segment * (chunksize * 32u)) / chunksize;
        const u32 nMasks = fullsegments + (additional > 0u ? 1u : 0u);
        u32* masks = on_page_masks(h, page, nMasks);
        rmw(h, &masks[segment], OP_AND, ~(1u << within));
        rmw(h, PTE_BITMASK(h, page), OP_AND, ~(1u << segment));
    }
    else
    {
        const u32 segment = inpage / chunksize;
        rmw(h, PTE_BITMASK(h, page), OP_AND, ~(1u << segment));
    }

    u32 oldfill;
    if(h->cfg.resetfreedpages)
    {
        oldfill = cas(h, PTE_COUNT(h, page), 1u, P);
        if(oldfill == 1u)
        {
            page_init(h, page);
            cas(h, PTE_CHUNKSIZE(h, page), chunksize, 0u);
            fence_device();
            rmw(h, PTE_COUNT(h, page), OP_SUB, P);
        }
        else
            oldfill = rmw(h, PTE_COUNT(h, page), OP_SUB, 1u);
    }
    else
        oldfill = rmw(h, PTE_COUNT(h, page), OP_SUB, 1u);

    if(oldfill == (MUT(RESET_WHEN_EMPTY) ? 1u : P / 2u / chunksize))
    {
        const u32 region = page / h->cfg.regionsize;
        if(!MUT(NO_REGION_RESET))
            rmw(h, &h->regions[region], OP_EXCH, 0u);
        const u32 pagesperblock = h->numpages / h->accessblocks;
        const u32 block = page / pagesperblock;
        if(!MUT(NO_BLOCK_MIN))
            rmw(h, &h->firstfreeblock, OP_MIN, block);
    }
}

/* markpages (Scatter.hpp:745-764) */
static int markpages(orc_heap* h, u32 startpage, u32 pages, u32 bytes)
{
    u32 abort_at = UINT32_MAX;
    for(u32 p = startpage; p < startpage + pages; ++p)
    {
        if(cas(h, PTE_CHUNKSIZE(h, p), 0u, bytes) != 0u)
        {
            abort_at = p;
            break;
        }
    }
    if(abort_at == UINT32_MAX)
        return 1;
    for(u32 p = startpage; p < abort_at; ++p)
        cas(h, PTE_CHUNKSIZE(h, p), bytes, 0u);
    return 0;
}

/* allocPageBasedSingleRegion (Scatter.hpp:775-813); page index numpages is
 * "not free" (divergence D2). */
static uint8_t* alloc_page_based_single_region(orc_heap* h, u32 startpage, u32 endpage, u32 bytes)
{
    const u32 pagestoalloc = divup32(bytes, h->cfg.pagesize);
    u32 freecount = 0;
    int left_free = 0;
    for(u32 sp = startpage + 1u; sp > endpage;)
    {
        --sp;
        const int is_free = sp < h->numpages && vload(PTE_CHUNKSIZE(h, sp)) == 0u;
        if(is_free)
        {
            if(++freecount == pagestoalloc)
            {
                if(markpages(h, sp, pagestoalloc, bytes))
                {
                    if(!left_free && !MUT(NO_FFPB_CAS))
                        cas(h, &h->firstFreePageBased, startpage, sp - 1u);
                    return page_data(h, sp);
                }
            }
        }
        else
        {
            left_free = 1;
            freecount = 0;
        }
    }
    return NULL;
}

/* allocPageBasedSingle (Scatter.hpp:824-840) */
static uint8_t* alloc_page_based_single(orc_heap* h, u32 bytes)
{
    while(rmw(h, &h->pagebasedMutex, OP_EXCH, 1u) != 0u)
    {
    }
    const u32 spage = vload(&h->firstFreePageBased);
    uint8_t* res = alloc_page_based_single_region(h, spage, 0u, bytes);
    if(res == NULL)
        res = alloc_page_based_single_region(h, h->numpages, spage, bytes);
    rmw(h, &h->pagebasedMutex, OP_EXCH, 0u);
    return res;
}

/* deallocPageBased (Scatter.hpp:874-886) */
static void dealloc_page_based(orc_heap* h, u32 page, u32 bytes)
{
    const u32 pages = divup32(bytes, h->cfg.pagesize);
    for(u32 p = page; p < page + pages; ++p)
        page_init(h, p);
    fence_device();
    for(u32 p = page; p < page + pages; ++p)
        cas(h, PTE_CHUNKSIZE(h, p), bytes, 0u);
    if(!MUT(NO_FFPB_MAX))
        rmw(h, &h->firstFreePageBased, OP_MAX, page + pages - 1u);
}

/* create (Scatter.hpp:895-912); allocPageBased (:848-866) reduces to one call
 * with a single active lane. */
static uint8_t* scatter_create(orc_heap* h, u32 bytes)
{
    if(bytes == 0u)
        return NULL;
    if(bytes <= h->cfg.pagesize)
        return alloc_chunked(h, bytes);
    return alloc_page_based_single(h, bytes);
}

/* destroy (Scatter.hpp:918-946), divergence D3 for chunksize 0. */
static void scatter_destroy(orc_heap* h, uint8_t* mem)
{
    if(mem == NULL)
        return;
    const u32 P = h->cfg.pagesize;
    const u32 page = (u32) ((u64) (mem - h->pool) / P);
    const u32 chunksize = vload(PTE_CHUNKSIZE(h, page));
    if(chunksize == 0u)
        return;
    const u32 inpage = (u32) (mem - page_data(h, page));
    const u32 block = inpage / chunksize;
    const u32 inblock = inpage - block * chunksize;
    if(inblock != 0u)
    {
        u32* counter = (u32*) (page_data(h, page) + (u64) block * chunksize);
        const u32 old = rmw(h, counter, OP_SUB, 1u);
        if(old != 1u)
            return;
        mem = (uint8_t*) counter;
    }
    if(chunksize < P)
        dealloc_chunked(h, mem, page, chunksize);
    else
        dealloc_page_based(h, page, chunksize);
}

/* ------------------------------------------------------------- geometry */
static int derive_constants(const orc_config* c, u32* HT, u32* minChunk, u32* maxMasks)
{
    const u32 P = c->pagesize;
    *HT = (P - 2u * 4u) / 33u;
    *minChunk = (P - 32u * 4u) / (32u * 32u);
    const u32 minSegment = 32u * *minChunk + 4u;
    const u32 onPage = *minChunk > *HT ? 0u : (P + (minSegment - 1u)) / minSegment;
    *maxMasks = umin(32u, onPage);
    return 0;
}

static const char* validate(const orc_config* c)
{
    if(c->pagesize < 128u || c->pagesize > (1u << 30))
        return "pagesize must be in [128, 2^30]";
    if(c->regionsize == 0u)
        return "regionsize must be > 0";
    if(c->accessblocksize == 0u)
        return "accessblocksize must be > 0";
    if(c->dataAlignment == 0u || (c->dataAlignment & (c->dataAlignment - 1u)) != 0u)
        return "dataAlignment must be a power of two";
    return NULL;
}

/* initDeviceFunction geometry (Scatter.hpp:962-1014), D1 for numblocks == 0. */
int orc_compute_geometry(const orc_config* c, uint64_t memsize, orc_geometry* g)
{
    if(validate(c) != NULL)
        return -1;
    memset(g, 0, sizeof(*g));
    derive_constants(c, &g->hierarchyThreshold, &g->minChunkSize, &g->maxOnPageMasks);
    const u64 P = c->pagesize;
    const u64 R = c->regionsize;
    u32 numregions = (u32) (memsize / (R * (12u + P) + 4u));
    u32 numpages = numregions * (u32) R;
    const u64 memorysize = (u64) numpages * P;
    u64 numblocks = memorysize / c->accessblocksize;
    if(numblocks == 0u)
    {
        numblocks = 1u;
        g->clampedBlocks = 1u;
    }
    const u64 memoryperblock = memorysize / numblocks;
    const u64 pagesperblock = memoryperblock / P;
    const u64 regionsperblock = pagesperblock / R;
    numregions = (u32) (numblocks * regionsperblock);
    numpages = numregions * (u32) R;
    u64 ptes = (u64) numpages * P;
    u64 regions = ptes + 12u * (u64) numpages;
    if(regions + 4u * (u64) numregions > memsize)
    {
        --numregions;
        numpages = umin(numregions * (u32) R, numpages);
        ptes = (u64) numpages * P;
        regions = ptes + 12u * (u64) numpages;
    }
    g->numpages = numpages;
    g->numregions = numregions;
    g->accessblocks = (u32) numblocks;
    g->pagesperblock = g->accessblocks ? numpages / g->accessblocks : 0u;
    g->ptesOffset = ptes;
    g->regionsOffset = regions;
    g->usedBytes = regions + 4u * (u64) numregions;
    return 0;
}

void orc_default_config(orc_config* c)
{
    memset(c, 0, sizeof(*c));
    c->pagesize = 4096u;
    c->regionsize = 16u;
    c->accessblocksize = 2ull * 1024u * 1024u * 1024u;
    c->wastefactor = 2u;
    c->resetfreedpages = 0u;
    c->hashingK = 38183u;
    c->hashingDistMP = 17497u;
    c->hashingDistWP = 1u;
    c->hashingDistWPRel = 1u;
    c->dataAlignment = 16u;
}

orc_heap* orc_create(const orc_config* cfg, uint64_t memsize, const char** err)
{
    const char* e = validate(cfg);
    orc_geometry g;
    if(e == NULL && orc_compute_geometry(cfg, memsize, &g) != 0)
        e = "invalid geometry";
    if(e == NULL && g.numpages == 0u)
        e = "heap too small: zero pages";
    if(e != NULL)
    {
        if(err)
            *err = e;
        return NULL;
    }
    orc_heap* h = (orc_heap*) calloc(1, sizeof(orc_heap));
    h->cfg = *cfg;
    h->HT = g.hierarchyThreshold;
    h->minChunkSize = g.minChunkSize;
    h->maxOnPageMasks = g.maxOnPageMasks;
    h->memsize = memsize;
    /* anonymous mapping: zero-filled and populated lazily */
    void* p = mmap(NULL, memsize, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if(p == MAP_FAILED)
    {
        free(h);
        if(err)
            *err = "mmap failed";
        return NULL;
    }
    h->pool = (uint8_t*) p;
    h->numpages = g.numpages;
    h->numregions = g.numregions;
    h->accessblocks = g.accessblocks;
    h->clamped = g.clampedBlocks;
    h->ptes = (u32*) (h->pool + g.ptesOffset);
    h->regions = (u32*) (h->pool + g.regionsOffset);
    /* init loop (Scatter.hpp:1026-1048); the mapping is already zero, so
     * only the page tails need an explicit store to match the reference's
     * write pattern. */
    for(u32 i = 0; i < h->numpages; ++i)
    {
        h->ptes[3u * (u64) i] = 0u;
        h->ptes[3u * (u64) i + 1u] = 0u;
        h->ptes[3u * (u64) i + 2u] = 0u;
        page_init(h, i);
    }
    for(u32 i = 0; i < h->numregions; ++i)
        h->regions[i] = 0u;
    h->firstfreeblock = 0u;
    h->pagebasedMutex = 0u;
    h->firstFreePageBased = h->numpages - 1u;
    h->atomic_mode = ORC_ATOMIC_LOCKFREE;
    return h;
}

void orc_destroy(orc_heap* h)
{
    if(!h)
        return;
    munmap(h->pool, h->memsize);
    free(h);
}

void orc_get_geometry(const orc_heap* h, orc_geometry* g)
{
    orc_compute_geometry(&h->cfg, h->memsize, g);
}

void orc_set_atomic_mode(orc_heap* h, int mode)
{
    h->atomic_mode = mode;
}

/* ------------------------------------------------------------ public API */
/* DeviceAllocator::malloc (device_allocator.hpp:72-82): pad (size_t -> u32),
 * Noop collect, create, isOOM -> ReturnNull, Noop distribute. */
uint64_t orc_malloc(orc_heap* h, uint64_t bytes)
{
    const u32 req = apply_padding(h, (u32) bytes);
    uint8_t* p = scatter_create(h, req);
    if(p == NULL)
        return ORC_NULL;
    return (u64) (p - h->pool);
}

void orc_free(orc_heap* h, uint64_t offset)
{
    if(offset == ORC_NULL)
        return;
    scatter_destroy(h, h->pool + offset);
}

/* countFreeChunksInPage (Scatter.hpp:1112-1139) */
static u32 count_free_chunks(orc_heap* h, u32 page, u32 chunksize)
{
    const u32 P = h->cfg.pagesize;
    const u32 filled = vload(PTE_COUNT(h, page));
    if(chunksize <= h->HT)
    {
        const u32 segmentsize = chunksize * 32u + 4u;
        const u32 fullsegments = umin(32u, P / segmentsize);
        const u32 additional = calc_additional(h, fullsegments, segmentsize, chunksize);
        return fullsegments * 32u + additional - filled;
    }
    return umin(P / chunksize, 32u) - filled;
}

/* getAvailaibleSlotsDeviceFunction (Scatter.hpp:1158-1219) over all pages;
 * the per-thread partial sums of the reference add up to this total. */
static u32 available_slots_raw(orc_heap* h, u64 slotSize)
{
    const u32 P = h->cfg.pagesize;
    u32 slots = 0u;
    if(slotSize < P)
    {
        const u32 maxchunksize = umin(P, h->cfg.wastefactor * (u32) slotSize);
        for(u32 p = 0; p < h->numpages; ++p)
        {
            u32 cs = vload(PTE_CHUNKSIZE(h, p));
            if(cs >= slotSize && cs <= maxchunksize)
                slots += count_free_chunks(h, p, cs);
            else if(cs == 0u)
            {
                cs = umax((u32) slotSize, apply_padding(h, h->minChunkSize));
                slots += count_free_chunks(h, p, cs);
            }
        }
    }
    else
    {
        const u32 pagestoalloc = divup32((u32) slotSize, P);
        u32 freecount = 0u;
        for(u32 p = h->numpages; p > 0u;)
        {
            --p;
            if(vload(PTE_CHUNKSIZE(h, p)) == 0u)
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

/* Allocator::getAvailableSlots pads first (allocator.hpp:227-233). */
uint32_t orc_available_slots(orc_heap* h, uint64_t slotsize)
{
    const u64 padded = apply_padding(h, (u32) slotsize);
    return available_slots_raw(h, padded);
}

void orc_run_stream(orc_heap* h, const int64_t* ops, uint64_t n, uint64_t* out)
{
    for(u64 i = 0; i < n; ++i)
    {
        if(ops[i] >= 0)
            out[i] = orc_malloc(h, (u64) ops[i]);
        else
        {
            const u64 idx = (u64) (-(ops[i] + 1));
            if(idx < i)
                orc_free(h, out[idx]);
            out[i] = 0;
        }
    }
}

void orc_read_pte(const orc_heap* h, uint32_t page, uint32_t* chunksize, uint32_t* count, uint32_t* bitmask)
{
    *chunksize = h->ptes[3u * (u64) page];
    *count = h->ptes[3u * (u64) page + 1u];
    *bitmask = h->ptes[3u * (u64) page + 2u];
}

uint32_t orc_read_region(const orc_heap* h, uint32_t region)
{
    return h->regions[region];
}

void orc_read_scalars(const orc_heap* h, uint32_t* firstFreePageBased, uint32_t* firstfreeblock)
{
    *firstFreePageBased = h->firstFreePageBased;
    *firstfreeblock = h->firstfreeblock;
}

const uint8_t* orc_pool(const orc_heap* h)
{
    return h->pool;
}

/* ------------------------------------------------------ CPU baseline run */
static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int orc_bench_concurrent(orc_heap* h, const uint32_t* sizes, uint64_t n, int nthreads, uint64_t* offsets,
                         orc_bench_result* res)
{
    memset(res, 0, sizeof(*res));
    if(nthreads <= 0)
        nthreads = omp_get_max_threads();
    res->threads_used = nthreads;
    u64 nulls = 0, bad = 0;

    double t0 = now_s();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : nulls)
    for(long long t = 0; t < (long long) n; ++t)
    {
        offsets[t] = orc_malloc(h, sizes[t]);
        if(offsets[t] == ORC_NULL && sizes[t] != 0u)
            ++nulls;
    }
    double t1 = now_s();
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : bad)
    for(long long t = 0; t < (long long) n; ++t)
    {
        if(offsets[t] == ORC_NULL)
            continue;
        u64* p = (u64*) (h->pool + offsets[t]);
        const u64 words = sizes[t] / 8u;
        for(u64 w = 0; w < words; ++w)
            p[w] = (u64) t;
    }
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : bad)
    for(long long t = 0; t < (long long) n; ++t)
    {
        if(offsets[t] == ORC_NULL)
            continue;
        const u64* p = (const u64*) (h->pool + offsets[t]);
        const u64 words = sizes[t] / 8u;
        for(u64 w = 0; w < words; ++w)
            bad += p[w] != (u64) t;
    }
    double t2 = now_s();
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for(long long t = 0; t < (long long) n; ++t)
        orc_free(h, offsets[t]);
    double t3 = now_s();
    res->t_malloc = t1 - t0;
    res->t_write_verify = t2 - t1;
    res->t_free = t3 - t2;
    res->nulls = nulls;
    res->mismatches = bad;
    return (int) (bad != 0);
}
