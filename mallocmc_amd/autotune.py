"""Autotuner of the batched (wave-cooperative) heap path.

Sweeps heap geometry (pagesize, regionsize) and wave tuning (aggregation
width, persistent launch, XCD slicing) on the caller's workload, measures
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
    ops_per_s: float = 0.0
    nulls: int = 0
    valid: bool = False
    seconds: float = 0.0

    def config(self, base: mmc.ScatterConfig) -> mmc.ScatterConfig:
        return dataclasses.replace(base, pagesize=self.pagesize, regionsize=self.regionsize)

    def tuning(self) -> mmc.Tuning:
        return mmc.Tuning(aggregation=self.aggregation, persistent=self.persistent, xcd_slices=self.xcd_slices)


DEFAULT_SPACE = dict(
    pagesize=(4096, 16384, 65536),
    regionsize=(8, 16, 32),
    aggregation=(1, 16, 64),
    persistent=(False, True),
    xcd_slices=(0, 8),
)


def measure(alloc: mmc.ScatterAllocator, size_batches, tuning: mmc.Tuning, rounds: int = 2):
    """(mallocs+frees)/s over `rounds` passes of every batch, invariant check on the last."""
    import torch
    n_ops = 0
    total_ms = 0.0
    nulls = 0
    ok = True
    for r in range(rounds):
        for sizes in size_batches:
            offs = torch.empty(sizes.numel(), dtype=torch.int64, device=sizes.device)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            alloc.wave_malloc_batch(sizes, offs, tuning=tuning)
            if r == rounds - 1:
                e1.record()
                torch.cuda.synchronize()
                total_ms += e0.elapsed_time(e1)
                alloc.write_pattern(offs, sizes, salt=7)
                res = alloc.verify_pattern(offs, sizes, salt=7)
                ok = ok and res["mismatches"] == 0 and res["misaligned"] == 0 and res["outside"] == 0
                nulls += res["nulls"]
                e0.record()
            alloc.wave_free_batch(offs, tuning=tuning)
            e1.record()
            torch.cuda.synchronize()
            if r == rounds - 1:
                total_ms += e0.elapsed_time(e1)
                n_ops += 2 * sizes.numel()
    return n_ops / (total_ms / 1e3), nulls, ok


def tune(base: mmc.ScatterConfig, heap_bytes: int, size_batches, space=None, log=None, max_nulls: int = 0):
    """Returns (best TunePoint, all points).  One heap per geometry serves all
    wave tunings of that geometry (free-all restores it between points when
    resetfreedpages is on; the restore is checked)."""
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
            for agg, pers, xcd in itertools.product(space["aggregation"], space["persistent"], space["xcd_slices"]):
                pt = TunePoint(pagesize, regionsize, agg, pers, xcd)
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
    best = max(valid, key=lambda p: p.ops_per_s)
    return best, points
