"""Autotuner of the batched (wave-cooperative) heap path.

Sweeps heap geometry (pagesize, regionsize) and wave tuning (aggregation
width, persistent launch, XCD slicing, requests per lane) on the caller's
workload, measures
mallocs+frees/s with HIP events on the launch stream, checks every candidate
against the verify_heap invariants, and returns the fastest valid point.
Candidates that return more nulls than the reference point, or break an
invariant, are rejected.
"""
from __future__ import annotations

import dataclasses
import itertools
import time

import mallocmc_amd as mmc

GiB = 1024 ** 3


@dataclasses.dataclass
class TunePoint:
    pagesize: int
    regionsize: int
    aggregation: int
    persistent: bool
    xcd_slices: int
    lane_requests: int = 1
    ops_per_s: float = 0.0
    nulls: int = 0
    valid: bool = False
    seconds: float = 0.0

    def config(self, base: mmc.ScatterConfig) -> mmc.ScatterConfig:
        return dataclasses.replace(base, pagesize=self.pagesize, regionsize=self.regionsize)

    def tuning(self) -> mmc.Tuning:
        return mmc.Tuning(aggregation=self.aggregation, persistent=self.persistent, xcd_slices=self.xcd_slices,
                          lane_requests=self.lane_requests)


DEFAULT_SPACE = dict(
    pagesize=(4096, 16384, 65536),
    regionsize=(8, 16, 32),
    aggregation=(1, 16, 64),
    persistent=(False, True),
    xcd_slices=(0, 8),
    lane_requests=(1, 2, 4),
)


def measure(alloc: mmc.ScatterAllocator, size_batches, tuning: mmc.Tuning, rounds: int = 2, timed: int = 1):
    """(mallocs+frees)/s: `rounds` passes over every batch, the last `timed`
    of them timed (the median pass rate is returned); the invariants are
    checked on the last pass."""
    import statistics

    import torch
    nulls = 0
    ok = True
    rates = []
    for r in range(rounds):
        n_ops = 0
        total_ms = 0.0
        for sizes in size_batches:
            offs = torch.empty(sizes.numel(), dtype=torch.int64, device=sizes.device)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            alloc.wave_malloc_batch(sizes, offs, tuning=tuning)
            e1.record()
            torch.cuda.synchronize()
            total_ms += e0.elapsed_time(e1)
            if r == rounds - 1:
                alloc.write_pattern(offs, sizes, salt=7)
                res = alloc.verify_pattern(offs, sizes, salt=7)
                ok = ok and res["mismatches"] == 0 and res["misaligned"] == 0 and res["outside"] == 0
                nulls += res["nulls"]
            e0.record()
            alloc.wave_free_batch(offs, tuning=tuning)
            e1.record()
            torch.cuda.synchronize()
            total_ms += e0.elapsed_time(e1)
            n_ops += 2 * sizes.numel()
        if r >= rounds - timed:
            rates.append(n_ops / (total_ms / 1e3))
    return statistics.median(rates), nulls, ok


def _same_point(a: TunePoint, b: TunePoint) -> bool:
    return (a.pagesize, a.regionsize, a.aggregation, bool(a.persistent), a.xcd_slices, a.lane_requests) == (
        b.pagesize, b.regionsize, b.aggregation, bool(b.persistent), b.xcd_slices, b.lane_requests)


def tune(base: mmc.ScatterConfig, heap_bytes: int, size_batches, space=None, log=None, max_nulls: int = 0,
         prefer: TunePoint | None = None, tie: float = 0.03, refine: int = 4, refine_passes: int = 5):
    """Returns (best TunePoint, all points).  One heap per geometry serves all
    wave tunings of that geometry (free-all restores it between points when
    resetfreedpages is on; the restore is checked).  Points within `tie` of
    the fastest valid point are taken as equal, since run-to-run noise is of
    that order; among them `prefer`, when given and measured, wins, so that
    repeated runs settle on one point.  The `refine` fastest points of the
    sweep (and `prefer`) are measured again over `refine_passes` timed passes
    each before that choice."""
    space = space or DEFAULT_SPACE
    points = []
    for pagesize, regionsize in itertools.product(space["pagesize"], space["regionsize"]):
        cfg = dataclasses.replace(base, pagesize=pagesize, regionsize=regionsize)
        try:
            alloc = mmc.ScatterAllocator(cfg, heap_bytes)
        except mmc.MallocMCError as e:
            if log:
                log(f"skip pagesize={pagesize} regionsize={regionsize}: {e}")
            continue
        try:
            slots0 = alloc.getAvailableSlots(64)
            for agg, pers, xcd, lr in itertools.product(space["aggregation"], space["persistent"],
                                                        space["xcd_slices"], space.get("lane_requests", (1,))):
                if pers and lr != 1:
                    continue  # persistent launches serve one request per lane
                pt = TunePoint(pagesize, regionsize, agg, pers, xcd, lr)
                t0 = time.perf_counter()
                pt.ops_per_s, pt.nulls, pt.valid = measure(alloc, size_batches, pt.tuning())
                pt.seconds = time.perf_counter() - t0
                restored = (not base.resetfreedpages) or alloc.getAvailableSlots(64) == slots0
                pt.valid = pt.valid and restored and pt.nulls <= max_nulls
                points.append(pt)
                if log:
                    log(f"tune {pt}")
        finally:
            alloc.close()
    valid = [p for p in points if p.valid]
    if not valid:
        raise mmc.MallocMCError("no valid tuning point")
    # one timed pass per point is noisy: re-measure the leaders (and the
    # preferred point) over `refine_passes` timed passes and rank by the median
    leaders = sorted(valid, key=lambda p: p.ops_per_s, reverse=True)[:refine]
    if prefer is not None:
        leaders += [p for p in valid if _same_point(p, prefer) and p not in leaders]
    for p in leaders:
        cfg = dataclasses.replace(base, pagesize=p.pagesize, regionsize=p.regionsize)
        alloc = mmc.ScatterAllocator(cfg, heap_bytes)
        try:
            ops, nulls, ok = measure(alloc, size_batches, p.tuning(), rounds=refine_passes + 1, timed=refine_passes)
        finally:
            alloc.close()
        p.ops_per_s, p.valid = ops, p.valid and ok and nulls <= max_nulls
        if log:
            log(f"tune refine {p}")
    leaders = [p for p in leaders if p.valid] or valid
    best = max(leaders, key=lambda p: p.ops_per_s)
    if prefer is not None:
        for p in leaders:
            if _same_point(p, prefer) and p.ops_per_s >= (1.0 - tie) * best.ops_per_s:
                return p, points
    return best, points
