"""Known-answer streams for the stateful heuristics of the Scatter walk.

Each stream below is a serialized request sequence on a tiny heap whose every
result -- the offset of each malloc, the region fill levels, the access-block
cursor (_firstfreeblock) and the multi-page cursor (_firstFreePageBased) at
checkpoints -- was derived by hand from the text of
src/include/mallocMC/creationPolicies/Scatter.hpp, reading it as the
reference's CPU backend executes it (warpSize 1, smid = warpid = laneid = 0,
mallocMC_utils.hpp:85-87).  The derivation is written beside each op.  These
vectors pin the oracle (tests/test_oracle_kats.py) and the GPU's serialized
mode (tests/test_gpu_kats.py) independently of each other.

Rules used throughout, with their lines:
  * The walk starts at page start_in_block + _firstfreeblock * pagesperblock
    (:535-538, :551).  start_in_block = (m * 38183 mod 2^32) mod pagesperblock
    with m = max(pad16(bytes), 16); every m used here (1024, 2048) makes
    m * 38183 a multiple of 1024, so the walk starts at the first page of the
    cursor's access block.
  * Pass 1 visits only regions whose fill level is below checklevel =
    regionsize * 3 / 4 (:553, :573-575); pass 2 visits every region
    (checklevel = regionsize + 1, :634).
  * A visited region whose pages all fail gets CAS(fill -> fill + 1) when
    fill + 1 <= regionsize (:612-617).
  * When the walk steps onto the first page of an access block with a larger
    index than the cursor, the cursor moves there (:620-629).
  * A flat page (chunk > hierarchy threshold (P-8)/33 = 123) holds
    min(P / chunk, 32) chunks and hands out its lowest free chunk (usespot
    from spot 0, :329-347); a full page's reservation fails and is rolled
    back (:494-517).
  * A free whose count before the decrement equals P / 2 / chunk sets the
    region's fill level to 0 and the cursor to min(cursor, block of the
    page) (:727-735).  With resetfreedpages the last free also returns the
    page to chunk size 0 (:686-716); every stream here reuses one chunk size
    per page, so both settings give the same results.
  * A multi-page request scans from _firstFreePageBased down to page 0 for
    the first run of free pages (:775-813, :831-835); when no used page lay
    above the run in that scan the cursor is CAS'd from the scan's start to
    run_start - 1, in uint32 arithmetic (:796-801); a multi-page free raises
    it with atomicMax(cursor, last page of the run) (:885).

Op encoding: M(bytes, want) mallocs and expects the pool offset `want`
(None = null); F(i) frees the result of op i.  The stream form (for
run_stream / mmc_serial_stream) is bytes for a malloc and -(i + 1) for a free.
"""
from __future__ import annotations

import dataclasses

U32_MAX = 2 ** 32 - 1
P = 4096


@dataclasses.dataclass(frozen=True)
class M:
    bytes: int
    want: int | None


@dataclasses.dataclass(frozen=True)
class F:
    op: int


@dataclasses.dataclass(frozen=True)
class State:
    regions: tuple
    firstfreeblock: int
    firstFreePageBased: int
    slots: tuple = ()  # (size, getAvailableSlots(size)) pairs


@dataclasses.dataclass(frozen=True)
class Geometry:
    """A tiny heap and its closed-form geometry (Scatter.hpp:962-1014):
    numregions = heap // (R * (12 + P) + 4), then redistributed over
    numpages * P // accessblocksize access blocks."""
    name: str
    regionsize: int
    heap: int
    accessblocksize: int
    numpages: int
    numregions: int
    accessblocks: int
    pagesperblock: int


# 32880 // (2 * 4108 + 4) = 4 regions of 2 pages; 8 * 4096 // 32768 = 1 block.
ONE_BLOCK_R2 = Geometry("one access block, regionsize 2", 2, 32880, 32768, 8, 4, 1, 8)
# 32872 // (4 * 4108 + 4) = 2 regions of 4 pages; 8 * 4096 // 16384 = 2 blocks
# of one region each.
TWO_BLOCKS_R4 = Geometry("two access blocks, regionsize 4", 4, 32872, 16384, 8, 2, 2, 4)


@dataclasses.dataclass(frozen=True)
class Kat:
    name: str
    geometry: Geometry
    ops: tuple
    states: dict  # op index -> State after that op
    mutants: tuple  # oracle mutant builds (scatter_oracle.c ORC_MUT_*) this stream must catch

    def stream(self):
        return [op.bytes if isinstance(op, M) else -(op.op + 1) for op in self.ops]

    def expected(self):
        """Per-op result as the serialized stream reports it: offset, -1 for
        null, 0 for a free."""
        return [(-1 if op.want is None else op.want) if isinstance(op, M) else 0 for op in self.ops]


def _fill(n, size, first_offset=0):
    return [M(size, first_offset + i * size) for i in range(n)]


KATS = [
    # ------------------------------------------------------------------
    Kat("fill level rises on a failed visit and gates pass 1", ONE_BLOCK_R2, tuple(
        # 0-3: pages 0 and 1 (region 0) take two 2048 B chunks each.
        _fill(4, 2048)
        + [
            # 4: region 0 visited, pages 0 and 1 full -> CAS fill 0 -> 1
            #    (:612-617); region 1, page 2 is formatted.
            M(2048, 8192),
            # 5: page 0 count 2 -> 1; old count 2 != 4096/2/2048 = 1, no reset.
            F(0),
            # 6: checklevel = 2*3/4 = 1 and region 0's fill is 1: skipped in
            #    pass 1 although its chunk 0 is free; page 2 chunk 1.
            M(2048, 10240),
            # 7: page 2 full, page 3 formatted.
            M(2048, 12288),
        ]),
        {4: State((1, 0, 0, 0), 0, 7), 5: State((1, 0, 0, 0), 0, 7), 7: State((1, 0, 0, 0), 0, 7)},
        mutants=(1, 9)),
    # ------------------------------------------------------------------
    Kat("second pass visits regions at or above checklevel", ONE_BLOCK_R2, tuple(
        _fill(4, 2048)
        + [
            M(2048, 8192),    # 4: region 0 full -> fill 1; page 2
            M(2048, 10240),   # 5
            M(2048, 12288),   # 6: page 3
            M(2048, 14336),   # 7
            M(2048, 16384),   # 8: region 0 skipped (1 >= 1); region 1 full -> fill 1; page 4
            M(2048, 18432),   # 9
            M(2048, 20480),   # 10: page 5
            M(2048, 22528),   # 11
            M(2048, 24576),   # 12: regions 0, 1 skipped; region 2 full -> fill 1; page 6
            M(2048, 26624),   # 13
            M(2048, 28672),   # 14: page 7
            M(2048, 30720),   # 15
            # 16: pass 1 skips regions 0-2 and finds region 3 full (fill
            #     0 -> 1); pass 2 (checklevel 3) visits all four, full, and
            #     raises each fill 1 -> 2 (2 <= regionsize): null.
            M(2048, None),
            # 17: page 2 chunk 1 freed; old count 2 != 1, no reset.
            F(5),
            # 18: pass 1 skips every region (fill 2 >= 1); pass 2: region 0
            #     full (fill + 1 = 3 > 2, no CAS), region 1: page 2 chunk 0 is
            #     held by op 4, so usespot moves to chunk 1.
            M(2048, 10240),
            # 19: full again; fill levels stay at the cap 2.
            M(2048, None),
        ]),
        {12: State((1, 1, 1, 0), 0, 7), 16: State((2, 2, 2, 2), 0, 7), 18: State((2, 2, 2, 2), 0, 7),
         19: State((2, 2, 2, 2), 0, 7)},
        mutants=(2, 8)),
    # ------------------------------------------------------------------
    Kat("region reset when a 2-chunk page drops from 1 to 0", ONE_BLOCK_R2, tuple(
        _fill(4, 2048)
        + [
            M(2048, 8192),    # 4: region 0 fill 1
            F(0),             # 5: old count 2 != 4096/2/2048 = 1: no reset
            M(2048, 10240),   # 6: region 0 skipped; page 2 chunk 1
            # 7: page 0 count 1 -> 0: old count 1 == 1 -> region 0 fill := 0,
            #    cursor := min(0, 0) (:727-735).
            F(1),
            M(2048, 0),       # 8: region 0 visited again: page 0 chunk 0
            M(2048, 2048),    # 9
        ]),
        {5: State((1, 0, 0, 0), 0, 7), 7: State((0, 0, 0, 0), 0, 7), 9: State((0, 0, 0, 0), 0, 7)},
        mutants=(3,)),
    # ------------------------------------------------------------------
    Kat("region reset at half of a 4-chunk page, not when it empties", ONE_BLOCK_R2, tuple(
        # 0-7: 1024 B chunks, 4 per page (min(4096/1024, 32)), pages 0 and 1.
        _fill(8, 1024)
        + [
            M(1024, 8192),    # 8: region 0 full -> fill 1; page 2 chunk 0
            F(0),             # 9: page 0 count 4 -> 3; old 4 != 4096/2/1024 = 2
            F(1),             # 10: count 3 -> 2; old 3 != 2
            M(1024, 9216),    # 11: region 0 skipped (fill 1); page 2 chunk 1
            F(2),             # 12: count 2 -> 1; old 2 == 2 -> region 0 fill := 0
            M(1024, 0),       # 13: page 0 holds only chunk 3: lowest free chunk 0
            M(1024, 1024),    # 14
        ]),
        {8: State((1, 0, 0, 0), 0, 7), 11: State((1, 0, 0, 0), 0, 7), 12: State((0, 0, 0, 0), 0, 7)},
        mutants=(3, 4)),
    # ------------------------------------------------------------------
    Kat("access-block cursor advances on the walk and falls back on a reset", TWO_BLOCKS_R4, tuple(
        # checklevel = 4*3/4 = 3; one region of 4 pages per access block.
        _fill(8, 2048)            # 0-7: block 0 (pages 0-3)
        + [
            # 8: region 0 full -> fill 1; the walk steps onto page 4, the
            #    first page of block 1 > cursor 0 -> cursor := 1 (:620-629);
            #    page 4 chunk 0.
            M(2048, 16384),
            F(0),             # 9: page 0 count 2 -> 1; no reset
            # 10: the walk now starts at 0 + 1*4 = page 4 (:551): page 4
            #     chunk 1, although region 0 (fill 1 < 3) has a free chunk.
            M(2048, 18432),
            # 11: page 0 count 1 -> 0: old 1 == 4096/2/2048 -> region 0 fill 0,
            #     cursor := min(1, block 0) = 0 (:733-734).
            F(1),
            M(2048, 0),       # 12: the walk starts at page 0 again
            M(2048, 2048),    # 13
        ]),
        {8: State((1, 0), 1, 7), 9: State((1, 0), 1, 7), 11: State((0, 0), 0, 7), 13: State((0, 0), 0, 7)},
        mutants=(5, 7)),
    # ------------------------------------------------------------------
    Kat("access-block cursor only moves up while the walk wraps", TWO_BLOCKS_R4, tuple(
        _fill(8, 2048)
        + [
            M(2048, 16384),   # 8: cursor := 1
            F(0),             # 9: page 0 chunk 0 free, no reset
            M(2048, 18432),   # 10: page 4 chunk 1
            M(2048, 20480),   # 11: page 5
            M(2048, 22528),   # 12
            M(2048, 24576),   # 13: page 6
            M(2048, 26624),   # 14
            M(2048, 28672),   # 15: page 7
            M(2048, 30720),   # 16
            # 17: walk from page 4: region 1 full -> fill 1; the walk wraps to
            #     page 0, block 0 is not > cursor 1 so the cursor stays 1
            #     (:627); region 0 (fill 1 < 3): page 0 chunk 0.
            M(2048, 0),
            F(2),             # 18: page 1 chunk 0 free (old count 2, no reset)
            # 19: walk from page 4 again: region 1 full -> fill 2; region 0:
            #     page 0 full, page 1 chunk 0.
            M(2048, 4096),
        ]),
        {17: State((1, 1), 1, 7), 19: State((1, 2), 1, 7)},
        mutants=(6,)),
    # ------------------------------------------------------------------
    Kat("_firstFreePageBased follows claimed and released runs", ONE_BLOCK_R2, (
        # 0: 2-page run: scan from 7: pages 7, 6 free -> run at 6; no used
        #    page above it -> CAS(cursor, 7 -> 6 - 1 = 5) (:796-801).
        M(8192, 24576),
        M(8192, 16384),   # 1: scan from 5: run at 4; cursor 5 -> 3
        # 2: pages 6, 7 released -> atomicMax(cursor, 6 + 2 - 1) = 7 (:885).
        F(0),
        M(8192, 24576),   # 3: scan from 7: run at 6 again; cursor 7 -> 5
        F(1),             # 4: pages 4, 5 released: max(5, 5) = 5
        M(8192, 16384),   # 5: scan from 5: run at 4; cursor 5 -> 3
    ), {0: State((0, 0, 0, 0), 0, 5), 1: State((0, 0, 0, 0), 0, 3), 2: State((0, 0, 0, 0), 0, 7),
        3: State((0, 0, 0, 0), 0, 5), 5: State((0, 0, 0, 0), 0, 3)},
        mutants=(10, 11)),
    # ------------------------------------------------------------------
    Kat("_firstFreePageBased wraps below page 0 and stops multi-page requests", ONE_BLOCK_R2, (
        M(8192, 24576),   # 0: cursor 7 -> 5
        M(8192, 16384),   # 1: cursor 5 -> 3
        M(8192, 8192),    # 2: cursor 3 -> 1
        # 3: scan from 1: pages 1, 0 free -> run at 0 with no used page above
        #    it: CAS(cursor, 1 -> 0 - 1), which wraps to 2^32 - 1 (uint32).
        M(8192, 0),
        # 4: pages 2, 3 released; atomicMax(2^32 - 1, 3) keeps 2^32 - 1.
        F(2),
        # 5: the first scan's loop starts at cursor + 1 = 0, which is not > 0,
        #    and the second starts at numpages + 1 = 9, not > 2^32 - 1 (:785,
        #    :831-835): no page is looked at, null, while pages 2 and 3 are
        #    free (getAvailableSlots counts that run, :1176-1196).
        M(8192, None),
        # 6: the chunked path is unaffected: pages 0, 1 carry chunk size 8192
        #    (not in [2048, 4096]), region 0 fill 0 -> 1; page 2 formatted.
        M(2048, 8192),
        M(8192, None),    # 7
    ), {3: State((0, 0, 0, 0), 0, U32_MAX), 4: State((0, 0, 0, 0), 0, U32_MAX),
        5: State((0, 0, 0, 0), 0, U32_MAX, slots=((8192, 1),)), 7: State((1, 0, 0, 0), 0, U32_MAX)},
        mutants=(10,)),
]


def check_geometry(g: Geometry, numpages, numregions, accessblocks, pagesperblock):
    got = (numpages, numregions, accessblocks, pagesperblock)
    want = (g.numpages, g.numregions, g.accessblocks, g.pagesperblock)
    assert got == want, f"{g.name}: geometry {got} != closed form {want}"


def run_on_oracle(orc, kat: Kat, resetfreedpages: bool, library=None) -> list:
    """Runs a stream op by op on the oracle (or a mutant build) and returns the
    differences from the hand-derived results."""
    g = kat.geometry
    cfg = orc.make_config(pagesize=P, regionsize=g.regionsize, accessblocksize=g.accessblocksize,
                          resetfreedpages=resetfreedpages)
    diffs = []
    with orc.OracleHeap(cfg, g.heap, library=library) as h:
        geo = h.geometry
        check_geometry(g, geo.numpages, geo.numregions, geo.accessblocks, geo.pagesperblock)
        results = []
        for i, op in enumerate(kat.ops):
            if isinstance(op, M):
                got = h.malloc(op.bytes)
                results.append(got)
                if got != op.want:
                    diffs.append(f"op {i}: malloc({op.bytes}) = {got}, want {op.want}")
            else:
                h.free(results[op.op])
                results.append(None)
            st = kat.states.get(i)
            if st is not None:
                regions = tuple(h.region(r) for r in range(geo.numregions))
                ffpb, ffb = h.scalars()
                if regions != st.regions:
                    diffs.append(f"after op {i}: regions {regions}, want {st.regions}")
                if ffb != st.firstfreeblock:
                    diffs.append(f"after op {i}: _firstfreeblock {ffb}, want {st.firstfreeblock}")
                if ffpb != st.firstFreePageBased:
                    diffs.append(f"after op {i}: _firstFreePageBased {ffpb}, want {st.firstFreePageBased}")
                for size, n in st.slots:
                    s = h.available_slots(size)
                    if s != n:
                        diffs.append(f"after op {i}: getAvailableSlots({size}) = {s}, want {n}")
    return diffs
