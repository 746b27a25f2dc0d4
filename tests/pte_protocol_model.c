/*
 * CPU model of the wave path's page claim and page reset on one page
 * (include/mallocMC/detail/wave_alloc.hpp: claimOnPage's reservation, claimBits'
 * atomicOr, backOff, resetEmpty and the free in WaveAlloc::free): the page's
 * paired 64-bit word (chunk size in the low half, count in the high half)
 * and its chunk mask.  Every step below is one atomic operation of the device
 * code; what a step computes from its result is local.
 *
 * The property: a lane that holds a chunk (its bit set by its own atomicOr
 * under a reservation) never sees the page's chunk size change until it
 * frees the chunk -- otherwise a reset and a new format could hand the same
 * bytes out again.
 *
 *   model   every interleaving of three processes (one freeing the page's
 *           last chunk, two claiming with chunk sizes and "fresh" flags of
 *           each combination), by depth-first search over the states; with
 *           the paired reservation (one 64-bit add that returns the chunk
 *           size of the same instant) no interleaving breaks the property;
 *           with the earlier split reservation (chunk-size CAS, then a 32-bit
 *           count add) some interleaving does, and the search prints it.
 *   threads 8 threads claim and free chunks of one page for a fixed number
 *           of rounds with the paired reservation, checking the property
 *           and that no chunk has two holders.
 * Prints "model paired: 0 violations" and so on; exit status 0 when the
 * paired protocol shows none and the split one shows at least one.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define C1 64u
#define C2 128u
#define CAP 2u /* chunks per page in the model, for either chunk size */

static uint32_t lo(uint64_t w) { return (uint32_t) w; }
static uint32_t hi(uint64_t w) { return (uint32_t) (w >> 32); }

/* ------------------------------------------------------------------ model */
enum
{
    PC_FORMAT,    /* fresh: CAS chunk size 0 -> cs (32-bit, low half) */
    PC_RESERVE,   /* paired: 64-bit add of 1 << 32; split: 32-bit count add */
    PC_FORMAT2,   /* reservation found the page unformatted: CAS 0 -> cs */
    PC_SNAP,      /* load the mask (a hint) */
    PC_OR,        /* atomicOr of the chosen bit */
    PC_FREE_AND,  /* free: clear the bit */
    PC_FREE_SUB,  /* free: count - 1 */
    PC_RESET_LD,  /* resetEmpty: load the pair */
    PC_RESET_CAS, /* resetEmpty: CAS (cs, 0) -> (0, 0) */
    PC_BACKOFF,   /* backOff: count - 1 */
    PC_DONE
};

typedef struct
{
    uint8_t pc, holding, bit, fresh;
    uint32_t cs, rcs, old, snap;
} Proc;

typedef struct
{
    uint64_t w;
    uint32_t mask;
    uint32_t pad;
    Proc p[3];
} State;

static int gSplit;          /* 1: the earlier split reservation */
static long gViolations, gStates;
static int gPrinted;

/* visited states: open addressing over full keys */
#define HSIZE (1u << 22)
static State* gTable;
static uint8_t* gUsed;

static uint64_t hashState(const State* s)
{
    const unsigned char* b = (const unsigned char*) s;
    uint64_t h = 1469598103934665603ull;
    for(size_t i = 0; i < sizeof(State); ++i)
        h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

static int visit(const State* s) /* 1 if new */
{
    uint64_t i = hashState(s) & (HSIZE - 1u);
    while(gUsed[i])
    {
        if(memcmp(&gTable[i], s, sizeof(State)) == 0)
            return 0;
        i = (i + 1u) & (HSIZE - 1u);
    }
    gUsed[i] = 1;
    gTable[i] = *s;
    ++gStates;
    return 1;
}

static uint32_t cas32lo(uint64_t* w, uint32_t expect, uint32_t val)
{
    const uint32_t cur = lo(*w);
    if(cur == expect)
        *w = ((uint64_t) hi(*w) << 32) | val;
    return cur;
}

/* the decision claimOnPage makes once it knows old and rcs */
static void decide(Proc* p)
{
    if(p->rcs != p->cs || p->old >= CAP)
        p->pc = PC_BACKOFF;
    else
        p->pc = PC_SNAP;
}

/* one atomic step of process i; returns 0 if it has none left */
static int step(State* s, int i)
{
    Proc* p = &s->p[i];
    switch(p->pc)
    {
    case PC_FORMAT:
        if(gSplit)
        {
            /* split: the chunk-size CAS is the check, issued first */
            const uint32_t c = cas32lo(&s->w, 0u, p->cs);
            p->rcs = c == 0u ? p->cs : c;
        }
        else
            (void) cas32lo(&s->w, 0u, p->cs);
        p->pc = PC_RESERVE;
        return 1;
    case PC_RESERVE:
        if(gSplit)
        {
            p->old = hi(s->w);
            s->w += 1ull << 32;
            decide(p);
        }
        else
        {
            const uint64_t pr = s->w;
            s->w += 1ull << 32;
            p->old = hi(pr);
            p->rcs = lo(pr);
            if(p->rcs == 0u)
                p->pc = PC_FORMAT2;
            else
                decide(p);
        }
        return 1;
    case PC_FORMAT2:
    {
        const uint32_t c = cas32lo(&s->w, 0u, p->cs);
        p->rcs = c == 0u ? p->cs : c;
        decide(p);
        return 1;
    }
    case PC_SNAP:
        p->snap = s->mask;
        p->pc = PC_OR;
        return 1;
    case PC_OR:
    {
        uint32_t b = 0u;
        while(b < CAP && ((p->snap >> b) & 1u))
            ++b;
        if(b == CAP)
        {
            /* the snapshot is full: read again (the reservation guarantees a bit) */
            p->pc = PC_SNAP;
            return 1;
        }
        const uint32_t o = s->mask;
        s->mask |= 1u << b;
        if(o & (1u << b))
        {
            p->snap = o | (1u << b);
            return 1; /* lost that bit: choose again from the OR's result */
        }
        p->bit = (uint8_t) b;
        p->holding = 1u;
        p->pc = PC_FREE_AND;
        return 1;
    }
    case PC_FREE_AND:
        s->mask &= ~(1u << p->bit);
        p->holding = 0u;
        p->pc = PC_FREE_SUB;
        return 1;
    case PC_FREE_SUB:
    case PC_BACKOFF:
    {
        const uint32_t o = hi(s->w);
        s->w -= 1ull << 32;
        p->pc = o == 1u ? PC_RESET_LD : PC_DONE;
        return 1;
    }
    case PC_RESET_LD:
    {
        const uint64_t pr = s->w;
        if(hi(pr) != 0u || lo(pr) == 0u)
            p->pc = PC_DONE;
        else
        {
            p->rcs = lo(pr);
            p->pc = PC_RESET_CAS;
        }
        return 1;
    }
    case PC_RESET_CAS:
        if(s->w == (uint64_t) p->rcs)
            s->w = 0u;
        p->pc = PC_DONE;
        return 1;
    default:
        return 0;
    }
}

static int broken(const State* s)
{
    for(int i = 0; i < 3; ++i)
        if(s->p[i].holding && lo(s->w) != s->p[i].cs)
            return 1;
    return 0;
}

static void search(const State* s)
{
    if(!visit(s))
        return;
    if(broken(s))
    {
        ++gViolations;
        if(!gPrinted)
        {
            gPrinted = 1;
            printf("  violation: page chunk size %u, count %u, mask %x; holders:", lo(s->w), hi(s->w), s->mask);
            for(int i = 0; i < 3; ++i)
                if(s->p[i].holding)
                    printf(" process %d (chunk size %u)", i, s->p[i].cs);
            printf("\n");
        }
        return;
    }
    int any = 0;
    for(int i = 0; i < 3; ++i)
    {
        State n = *s;
        if(step(&n, i))
        {
            any = 1;
            search(&n);
        }
    }
    if(!any && (s->mask != 0u || hi(s->w) != 0u))
    {
        ++gViolations; /* ended with a count or a bit left behind */
        if(!gPrinted)
        {
            gPrinted = 1;
            printf("  violation: final count %u, mask %x\n", hi(s->w), s->mask);
        }
    }
}

static long model(int split)
{
    gSplit = split;
    gViolations = 0;
    gPrinted = 0;
    long states = 0;
    const uint32_t sizes[2] = {C1, C2};
    for(int a = 0; a < 4; ++a)
        for(int b = a; b < 4; ++b)
        {
            memset(gUsed, 0, HSIZE);
            gStates = 0;
            State s;
            memset(&s, 0, sizeof(s));
            /* the page: formatted for C1, one chunk (bit 0) held by process 0 */
            s.w = (1ull << 32) | C1;
            s.mask = 1u;
            s.p[0].pc = PC_FREE_AND;
            s.p[0].holding = 1u;
            s.p[0].bit = 0u;
            s.p[0].cs = C1;
            const int ab[2] = {a, b};
            for(int j = 0; j < 2; ++j)
            {
                Proc* p = &s.p[1 + j];
                p->cs = sizes[ab[j] & 1];
                p->fresh = (uint8_t) (ab[j] >> 1);
                /* the split claim always issues its chunk-size CAS first; the
                   paired one only for a page its scan saw unformatted */
                p->pc = (split || p->fresh) ? PC_FORMAT : PC_RESERVE;
            }
            search(&s);
            states += gStates;
        }
    printf("model %s: %ld violations over %ld states\n", split ? "split" : "paired", gViolations, states);
    return gViolations;
}

/* ---------------------------------------------------------------- threads */
#define THREADS 8
#define ROUNDS 200000

static _Atomic uint64_t tW;
static _Atomic uint32_t tMask;
static _Atomic uint32_t tOwner[CAP];
static _Atomic long tViolations, tClaims;

static uint64_t mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint32_t tcas32lo(uint32_t expect, uint32_t val)
{
    uint64_t cur = atomic_load(&tW);
    for(;;)
    {
        if(lo(cur) != expect)
            return lo(cur);
        const uint64_t next = ((uint64_t) hi(cur) << 32) | val;
        if(atomic_compare_exchange_weak(&tW, &cur, next))
            return expect;
    }
}

static void tReleaseCount(void)
{
    const uint64_t pr = atomic_fetch_sub(&tW, 1ull << 32);
    if(hi(pr) == 1u)
    {
        uint64_t cur = atomic_load(&tW);
        if(hi(cur) == 0u && lo(cur) != 0u)
        {
            uint64_t expect = cur;
            atomic_compare_exchange_strong(&tW, &expect, 0ull);
        }
    }
}

static void* worker(void* arg)
{
    const uint32_t id = (uint32_t) (uintptr_t) arg + 1u;
    for(uint64_t r = 0; r < ROUNDS; ++r)
    {
        const uint64_t z = mix(((uint64_t) id << 40) | r);
        const uint32_t cs = (z & 1u) ? C2 : C1;
        const uint64_t seen = atomic_load(&tW);
        if(lo(seen) == 0u)
            (void) tcas32lo(0u, cs); /* fresh: format first */
        else if(lo(seen) != cs)
            continue;
        const uint64_t pr = atomic_fetch_add(&tW, 1ull << 32);
        uint32_t rcs = lo(pr);
        const uint32_t old = hi(pr);
        if(rcs == 0u)
        {
            const uint32_t c = tcas32lo(0u, cs);
            rcs = c == 0u ? cs : c;
        }
        if(rcs != cs || old >= CAP)
        {
            tReleaseCount();
            continue;
        }
        uint32_t bit = CAP;
        for(;;)
        {
            const uint32_t snap = atomic_load(&tMask);
            uint32_t b = 0u;
            while(b < CAP && ((snap >> b) & 1u))
                ++b;
            if(b == CAP)
                continue;
            const uint32_t o = atomic_fetch_or(&tMask, 1u << b);
            if(!(o & (1u << b)))
            {
                bit = b;
                break;
            }
        }
        atomic_fetch_add(&tClaims, 1);
        uint32_t none = 0u;
        if(!atomic_compare_exchange_strong(&tOwner[bit], &none, id))
            atomic_fetch_add(&tViolations, 1); /* two holders of one chunk */
        for(int k = 0; k < 4; ++k)
            if(lo(atomic_load(&tW)) != cs)
                atomic_fetch_add(&tViolations, 1); /* chunk size changed under a holder */
        uint32_t mine = id;
        if(!atomic_compare_exchange_strong(&tOwner[bit], &mine, 0u))
            atomic_fetch_add(&tViolations, 1);
        atomic_fetch_and(&tMask, ~(1u << bit));
        tReleaseCount();
    }
    return NULL;
}

static long threads(void)
{
    pthread_t t[THREADS];
    for(uintptr_t i = 0; i < THREADS; ++i)
        if(pthread_create(&t[i], NULL, worker, (void*) i) != 0)
        {
            printf("pthread_create failed\n");
            return -1;
        }
    for(int i = 0; i < THREADS; ++i)
        pthread_join(t[i], NULL);
    const uint64_t w = atomic_load(&tW);
    long v = atomic_load(&tViolations);
    if(hi(w) != 0u || atomic_load(&tMask) != 0u)
        ++v;
    printf("threads paired: %ld violations over %ld claims (final count %u, mask %x)\n", v, atomic_load(&tClaims),
           hi(w), atomic_load(&tMask));
    return v;
}

int main(int argc, char** argv)
{
    const char* what = argc > 1 ? argv[1] : "all";
    int ok = 1;
    if(strcmp(what, "threads") != 0)
    {
        gTable = malloc(sizeof(State) * HSIZE);
        gUsed = malloc(HSIZE);
        if(!gTable || !gUsed)
            return 2;
        ok = ok && model(0) == 0;
        ok = ok && model(1) > 0;
        free(gTable);
        free(gUsed);
    }
    if(strcmp(what, "model") != 0)
        ok = ok && threads() == 0;
    printf("%s\n", ok ? "pte protocol OK" : "pte protocol FAILED");
    return ok ? 0 : 1;
}
