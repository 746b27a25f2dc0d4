"""Pins the CPU oracle (oracle/scatter_oracle.c) before it is trusted.

The reference publishes no golden vectors and its CPU path cannot be built
here (alpaka needs Boost, absent from the image), so the oracle is pinned by:
  * closed forms of the geometry (Scatter.hpp:962-1014), the page capacities
    (Scatter.hpp:365-383, 494-510) and the first serialized offset of each
    size on a fresh heap (Scatter.hpp:535-551 with smid = warpid = 0);
  * the slot counts a probe build of the reference printed for 16,384 x
    malloc(16) on a 64 MiB heap at pagesize 4 KiB
    (4,031,488 -> 4,015,104 -> 4,031,488; first offset 458,752);
  * per-step invariants on brute-force random streams over tiny heaps, which
    hold whatever the bit tricks inside: no overlap, alignment, the
    wastefactor rule, lowest-free-chunk claim order, count == live chunks,
    and null only when no page can take the request;
  * special cases: 0 bytes, multi-page first fit from the top, and the
    pagesize-sized chunk that leaks one slot (Scatter.hpp:906-908, 942-945).
"""
import math

import numpy as np
import pytest

from mallocmc_amd import workloads as wl

MiB = 1024 ** 2
GiB = 1024 ** 3


def pad(b, a=16):
    return (b + a - 1) // a * a


def capacity(P, c):
    """Chunks per page (tryUsePage, Scatter.hpp:494-510), from the formulas."""
    HT = (P - 8) // 33
    if c <= HT:
        seg = 32 * c + 4
        full = min(32, P // seg)
        add = 0 if full == 32 else min(31, max(0, P - full * seg - 4) // c)
        return full * 32 + add
    return min(P // c, 32)


# --------------------------------------------------------------- geometry
GEOMETRY_TABLE = [
    # (P, AB, R, heap) -> pages, regions, access blocks, pages per block  (closed forms of Scatter.hpp:962-1014)
    ((4096, 2 * MiB, 16, 64 * MiB), (15872, 992, 31, 512)),
    ((16384, 2 * GiB, 16, 4 * GiB), (261936, 16371, 1, 261936)),
    ((16384, 2 * MiB, 8, 256 * MiB), (16256, 2032, 127, 128)),
    ((65536, 64 * MiB, 32, 256 * MiB), (4032, 126, 3, 1344)),
    ((65536, 2 * GiB, 16, 64 * GiB), (1048048, 65503, 31, 33808)),
    ((4096, 2 * GiB, 16, 32 * GiB), (8363520, 522720, 15, 557568)),
]


@pytest.mark.parametrize("args,want", GEOMETRY_TABLE)
def test_geometry_closed_forms(oracle_mod, args, want):
    P, AB, R, heap = args
    g = oracle_mod.geometry(oracle_mod.make_config(pagesize=P, accessblocksize=AB, regionsize=R), heap)
    assert (g.numpages, g.numregions, g.accessblocks, g.pagesperblock) == want
    # layout: [pages][12 B PTE per page][4 B per region] fits the heap
    assert g.ptesOffset == g.numpages * P
    assert g.regionsOffset == g.ptesOffset + 12 * g.numpages
    assert g.usedBytes <= heap
    assert g.numpages % g.accessblocks == 0


def test_geometry_constants(oracle_mod):
    for P, HT, mc in ((4096, 123, 3), (16384, 496, 15), (65536, 1985, 63)):
        g = oracle_mod.geometry(oracle_mod.make_config(pagesize=P), 256 * MiB)
        assert (g.hierarchyThreshold, g.minChunkSize, g.maxOnPageMasks) == (HT, mc, 32)


def test_heap_smaller_than_accessblock_is_clamped(oracle_mod):
    g = oracle_mod.geometry(oracle_mod.make_config(pagesize=4096, accessblocksize=2 * GiB), 1 * GiB)
    assert g.clampedBlocks == 1 and g.accessblocks == 1
    assert g.numpages == (1 * GiB // (16 * (12 + 4096) + 4)) * 16


# ------------------------------------------------------ capacity and slots
CAPACITY_TABLE = {
    4096: {16: 254, 32: 127, 48: 85, 64: 63, 96: 42, 112: 36, 128: 32, 256: 16, 512: 8, 1024: 4, 2048: 2, 4096: 1},
    16384: {16: 1016, 32: 510, 64: 255, 128: 127, 256: 63, 512: 32, 4096: 4},
    65536: {64: 1022, 128: 511, 256: 255, 512: 127, 1024: 63, 2048: 32, 4096: 16},
}


def test_capacity_formula_matches_table():
    for P, tab in CAPACITY_TABLE.items():
        for c, n in tab.items():
            assert capacity(P, c) == n, (P, c)


def test_remainder_never_negative_for_aligned_chunks():
    """calcAdditionalChunks' max(0, P - full*(32c+4) - 4) is never clamped for
    16-aligned chunk sizes at the configured pagesizes, so its int/unsigned
    reading cannot matter (Scatter.hpp:373-380)."""
    for P in (4096, 16384, 65536):
        HT = (P - 8) // 33
        for c in range(16, HT + 1, 16):
            seg = 32 * c + 4
            full = min(32, P // seg)
            if full != 32:
                assert P - full * seg - 4 >= 0, (P, c)


@pytest.mark.parametrize("P", [4096, 16384, 65536])
def test_empty_heap_slots_closed_form(oracle_mod, P):
    heap = 64 * MiB if P < 65536 else 256 * MiB
    with oracle_mod.OracleHeap(oracle_mod.make_config(pagesize=P, accessblocksize=8 * MiB), heap) as h:
        n = h.geometry.numpages
        mc = pad((P - 128) // 1024)
        for s in (1, 16, 17, 100, 500, 1000, P // 2, P - 1):
            assert h.available_slots(s) == n * capacity(P, max(pad(s), mc)), s
        for k in (1, 2, 3, 7):
            assert h.available_slots(k * P) == n // k


def test_probe_values_16B_64MiB(oracle_mod):
    """Values printed by a probe build of the reference on its CPU backend."""
    cfg = oracle_mod.make_config(pagesize=4096, accessblocksize=2 * MiB, resetfreedpages=True)
    with oracle_mod.OracleHeap(cfg, 64 * MiB) as h:
        assert h.available_slots(16) == 4031488
        offs = [h.malloc(16) for _ in range(16384)]
        assert offs[0] == 458752
        assert h.available_slots(16) == 4015104
        for o in offs:
            h.free(o)
        assert h.available_slots(16) == 4031488


# ---------------------------------------------------- first offset closed form
@pytest.mark.parametrize("P,AB,heap", [(4096, 2 * MiB, 64 * MiB), (16384, 2 * GiB, 512 * MiB),
                                       (65536, 64 * MiB, 256 * MiB), (1024, 1 * MiB, 16 * MiB)])
def test_first_offset_closed_form(oracle_mod, P, AB, heap):
    cfg = oracle_mod.make_config(pagesize=P, accessblocksize=AB)
    g = oracle_mod.geometry(cfg, heap)
    mc = pad((P - 128) // 1024)
    for b in (1, 16, 24, 64, 100, 496, 1000, P // 4, P // 2, P - 16, P):
        with oracle_mod.OracleHeap(cfg, heap) as h:
            m = max(pad(b), mc)
            want = ((m * 38183) % (2 ** 32) % g.pagesperblock) * P
            assert h.malloc(b) == want, b
    for k in (2, 3, 5):
        with oracle_mod.OracleHeap(cfg, heap) as h:
            assert h.malloc(k * P - 8) == (g.numpages - k) * P
    with oracle_mod.OracleHeap(cfg, heap) as h:
        assert h.malloc(0) is None


# --------------------------------------------------- brute-force invariants
def _check_state(h, P, live, nb, WF=2, mc=16):
    """Invariants over the whole heap after each op (live: offset -> padded size)."""
    per_page = {}
    for off, sz in live.items():
        per_page.setdefault(off // P, []).append((off, sz))
    for page, items in per_page.items():
        cs, count, _ = h.pte(page)
        if cs > P:  # multi-page allocation
            assert len(items) == 1 and items[0][0] == page * P
            continue
        assert cs > 0
        idx = set()
        for off, sz in items:
            assert (off - page * P) % cs == 0
            assert sz <= cs <= max(WF * sz, mc) or cs == sz
            i = (off - page * P) // cs
            assert i < capacity(P, cs)
            idx.add(i)
        assert len(idx) == len(items)
        assert count >= len(items)


def _longest_free_run(free):
    best = cur = 0
    for f in free:
        cur = cur + 1 if f else 0
        best = max(best, cur)
    return best


@pytest.mark.parametrize("seed,P,R,reset", [(11, 4096, 4, True), (12, 4096, 2, False), (13, 16384, 4, True),
                                            (14, 1024, 2, True)])
def test_bruteforce_invariants(oracle_mod, seed, P, R, reset):
    rng = np.random.default_rng(seed)
    heap = 8 * (12 + P) * R * 2 + 4096
    cfg = oracle_mod.make_config(pagesize=P, accessblocksize=heap // 2, regionsize=R, resetfreedpages=reset)
    mc = pad((P - 128) // 1024)
    cand = [16, 32, 48, 64, 100, 200, 500, P // 4, P // 2, P - 16, 2 * P]
    with oracle_mod.OracleHeap(cfg, heap) as h:
        g = h.geometry
        live = {}
        handles = []
        nulls = wrapped = 0
        for step in range(1500):
            if handles and rng.random() < 0.4:
                j = int(rng.integers(len(handles)))
                off = handles.pop(j)
                h.free(off)
                del live[off]
            else:
                b = int(rng.choice(cand))
                pb = pad(b)
                free_before = [h.pte(p)[0] == 0 for p in range(g.numpages)]
                ffpb_before = h.scalars()[0]
                off = h.malloc(b)
                if off is None:
                    nulls += 1
                    if pb <= P:
                        # null => no page can take the request
                        for p in range(g.numpages):
                            cs, cnt, _ = h.pte(p)
                            assert cs != 0
                            if pb <= cs <= min(P, max(2 * pb, mc)) and cs <= P:
                                assert cnt >= capacity(P, cs)
                    elif ffpb_before < g.numpages:
                        # null => no run of ceil(pb/P) free pages anywhere
                        assert _longest_free_run(free_before) < math.ceil(pb / P)
                    else:
                        # a run claimed at page 0 wrapped the cursor to 2^32-1,
                        # after which the page-based scans look at no page
                        # (Scatter.hpp:785, 796-801, 831-835)
                        wrapped += 1
                        assert ffpb_before == 2 ** 32 - 1
                    continue
                assert off % 16 == 0 and off + pb <= g.numpages * P
                for o, s in live.items():
                    assert off + pb <= o or o + s <= off, "overlap"
                if pb <= P:
                    page = off // P
                    cs = h.pte(page)[0]
                    i = (off - page * P) // cs
                    used = {(o - page * P) // cs for o in live if o // P == page}
                    assert i == min(set(range(capacity(P, cs))) - used), "lowest free chunk"
                live[off] = pb
                handles.append(off)
            _check_state(h, P, live, g.numpages, mc=mc)
            # Every page above _firstFreePageBased belongs to a live multi-page
            # run: the cursor only drops across the run a scan just claimed
            # (:796-801) and a released run raises it to its last page (:885).
            # So in a serialized stream the second scan (:835) never finds a run.
            ffpb = h.scalars()[0]
            for p in range(ffpb + 1, g.numpages):
                assert h.pte(p)[0] > P, (step, p, ffpb)
        assert nulls > 0, "streams are sized to reach OOM"
        assert nulls > wrapped


def test_pagesize_chunk_leaks_one_slot(oracle_mod):
    """A request of exactly pagesize is served chunked but released page-based,
    leaving count and bitmask set on its page (Scatter.hpp:874-886)."""
    P = 4096
    cfg = oracle_mod.make_config(pagesize=P, accessblocksize=8 * MiB, resetfreedpages=True)
    with oracle_mod.OracleHeap(cfg, 8 * MiB) as h:
        n = h.geometry.numpages
        s0 = h.available_slots(2048)
        off = h.malloc(P)
        page = off // P
        assert h.pte(page) == (P, 1, 1)
        h.free(off)
        assert h.pte(page) == (0, 1, 1)
        assert h.available_slots(2048) == s0 - 1
        assert h.available_slots(P) == n


def test_stream_api_matches_calls(oracle_mod):
    rng = np.random.default_rng(3)
    sizes = rng.choice([16, 64, 1000, 4096, 9000], size=500)
    ops = wl.random_stream(rng, 400, sizes)
    cfg = oracle_mod.make_config(pagesize=4096, accessblocksize=1 * MiB)
    with oracle_mod.OracleHeap(cfg, 4 * MiB) as a, oracle_mod.OracleHeap(cfg, 4 * MiB) as b:
        got = a.run_stream(ops)
        res = []
        for op in ops:
            if op >= 0:
                o = b.malloc(int(op))
                res.append(oracle_mod.ORC_NULL if o is None else o)
            else:
                prev = res[-op - 1]
                b.free(None if prev == oracle_mod.ORC_NULL else prev)
                res.append(0)
        np.testing.assert_array_equal(got, np.array(res, dtype=np.uint64))


def test_concurrent_baseline_invariants(oracle_mod):
    """The OpenMP CPU baseline keeps the verify_heap invariants with both atomic
    back-ends.  The reference's reset protocol (Scatter.hpp:686-716) can leave a
    fully freed page formatted when two frees race past CAS(count, 1 -> P), so
    after a mixed-size free-all only "no live chunk" (every count 0) is
    guaranteed; with one size the slot count is restored as well."""
    for sizes, restore in ((wl.mixed_sizes(1, 20000), False), (np.full(20000, 16, np.uint32), True)):
        for mode in (oracle_mod.ATOMIC_LOCKFREE, oracle_mod.ATOMIC_MUTEX16):
            cfg = oracle_mod.make_config(pagesize=16384, accessblocksize=2 * GiB, resetfreedpages=True)
            with oracle_mod.OracleHeap(cfg, 128 * MiB) as h:
                h.set_atomic_mode(mode)
                s0 = h.available_slots(16)
                res, offs = h.bench_concurrent(sizes, nthreads=4)
                assert res.mismatches == 0 and res.nulls == 0
                o = offs.astype(np.int64)
                order = np.argsort(o)
                ends = o[order] + sizes[order].astype(np.int64)
                assert (ends[:-1] <= o[order][1:]).all()
                assert all(h.pte(p)[1] == 0 for p in range(h.geometry.numpages))
                if restore:
                    assert h.available_slots(16) == s0


def test_first_null_fill_of_reference_protocol(oracle_mod):
    """Config-4 size mix on a 128 MiB heap, pagesize 64 KiB: the reference's
    protocol returns its first null only once no page is free (null only when
    no page can take the request), and its payload fill at that point stays
    well below 95%: a request above half a page takes a whole page (one
    chunk per page, Scatter.hpp:494-510)."""
    from oom_fill import oracle_first_null_fill
    res = oracle_first_null_fill(oracle_mod, 128 * 1024 ** 2)
    for r in res:
        assert r["free_pages"] == 0, r
        assert 0.55 < r["fill_fraction"] < 0.8, r
