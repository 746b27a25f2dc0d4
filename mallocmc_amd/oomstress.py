"""Fragmentation / OOM stress.

Heap 64 GiB at pagesize 64 KiB.  2,097,152 lanes allocate bimodal sizes (80%
uniform 16..256 B, 20% uniform 8..60 KiB) until the first null any of them
sees (default), or until the heap returns null for each of them
(first_null=False: verify_heap's AllocAll loop, tests/verify_heap.cpp:333-361
of the reference, where every lane pays a failing full-heap scan), then a
seeded random half of all allocations is freed and the lanes refill; 5 cycles.  After every fill: no two live allocations overlap,
every allocation holds its pattern and is aligned and inside the page area;
the fill fraction is reported.  After the last cycle everything is freed and
getAvailableSlots must return to its initial values.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

import mallocmc_amd as mmc

GiB = 1024 ** 3
MiB = 1024 ** 2


class MmcFillDesc(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("lanes", ctypes.c_uint32),
        ("p_small_permille", ctypes.c_uint32),
        ("small_lo", ctypes.c_uint32),
        ("small_hi", ctypes.c_uint32),
        ("large_lo", ctypes.c_uint32),
        ("large_hi", ctypes.c_uint32),
        ("cap", ctypes.c_uint64),
        ("d_offsets", ctypes.c_void_p),
        ("d_sizes", ctypes.c_void_p),
        ("d_count", ctypes.c_void_p),
        ("d_stop", ctypes.c_void_p),
    ]


def _bind():
    L = mmc.lib()
    if not getattr(L, "_fill_bound", False):
        L.mmc_wave_fill_until_null.argtypes = [ctypes.c_void_p, ctypes.POINTER(MmcFillDesc),
                                               ctypes.POINTER(mmc.MmcTuning), ctypes.c_void_p]
        L.mmc_wave_fill_until_null.restype = ctypes.c_int
        L._fill_bound = True
    return L


def _no_overlap(torch, offs, sizes):
    o = offs
    s = ((sizes.to(torch.int64) + 15) // 16) * 16
    order = torch.argsort(o)
    o, s = o[order], s[order]
    if o.numel() < 2:
        return True
    return int(((o[:-1] + s[:-1]) > o[1:]).sum().item()) == 0


def run_default(heap_bytes: int = 64 * GiB, lanes: int = 2 * 1024 * 1024, cycles: int = 5, pagesize: int = 65536,
                regionsize: int = 32, tuning: mmc.Tuning | None = None, log=None, cap: int | None = None,
                first_null: bool = True, verify_snapshots: bool = False, accessblocksize: int | None = None) -> dict:
    """first_null: every lane stops at the first null any lane sees (the
    heap's fill at its first failure); else each lane allocates until its own
    null (verify_heap's AllocAll).  verify_snapshots: after every fill, also
    check a snapshot of the heap's metadata against the complete list of live
    allocations (mallocmc_amd/snapshot.py).

    Defaults: one access block over the whole heap (`accessblocksize` None),
    regionsize 32 and one size class per octave (class_split 0).  Every walk
    starts at its hashed page within the access block under the cursor; with
    2 GiB blocks all of them start in the same 32768 pages and meet at the
    same last free pages as the heap fills.  Measured on the 64 GiB heap, 4
    cycles each, refill: one access block 0.024-0.028 s, 8 GiB blocks
    0.19-0.21 s, 2 GiB blocks 0.21-0.27 s (regionsize 32; at 2 GiB blocks
    regionsize 16 with four classes per octave 0.29-0.34 s).  The fill at the
    first null stays 0.68-0.71 at all of them, with no page left free."""
    import torch
    L = _bind()
    tuning = tuning or mmc.Tuning(aggregation=64, class_split=0)
    cfg = mmc.ScatterConfig(pagesize=pagesize, accessblocksize=accessblocksize or heap_bytes, regionsize=regionsize,
                            resetfreedpages=True, flags=mmc.FLAG_FIX_PAGESIZE_LEAK)
    alloc = mmc.ScatterAllocator(cfg, heap_bytes)
    g = alloc.geometry()
    page_bytes = int(g.numpages) * pagesize
    probe = (16, 256, 8192, pagesize, 4 * pagesize)
    slots0 = [alloc.getAvailableSlots(s) for s in probe]
    cap = cap or max(1 << 20, heap_bytes // 4096)
    offs = torch.full((cap,), -1, dtype=torch.int64, device="cuda")
    sizes = torch.zeros(cap, dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    stop = torch.zeros(1, dtype=torch.int32, device="cuda")
    rng = np.random.default_rng(4)
    results = []
    ok = True
    for cyc in range(cycles):
        d = MmcFillDesc(seed=1000 + cyc, lanes=lanes, p_small_permille=800, small_lo=16, small_hi=256,
                        large_lo=8 * 1024, large_hi=60 * 1024, cap=cap, d_offsets=offs.data_ptr(),
                        d_sizes=sizes.data_ptr(), d_count=count.data_ptr(),
                        d_stop=stop.data_ptr() if first_null else None)
        stop.zero_()
        before = int(count.item())
        t0 = time.perf_counter()
        mmc._check(L.mmc_wave_fill_until_null(alloc.handle, ctypes.byref(d), ctypes.byref(tuning.to_c()),
                                              mmc._stream_ptr()), "fill")
        torch.cuda.synchronize()
        t_fill = time.perf_counter() - t0
        n = min(int(count.item()), cap)
        live_o, live_s = offs[:n], sizes[:n]
        overlap_ok = _no_overlap(torch, live_o, live_s)
        alloc.write_pattern(live_o, live_s, salt=cyc << 48)
        v = alloc.verify_pattern(live_o, live_s, salt=cyc << 48)
        cyc_ok = overlap_ok and v["mismatches"] == 0 and v["misaligned"] == 0 and v["outside"] == 0 and v["nulls"] == 0
        snap_report = None
        if verify_snapshots:
            from mallocmc_amd import snapshot
            snap_report = snapshot.verify(snapshot.take(alloc), live_o, live_s, complete=True)
            cyc_ok = cyc_ok and snap_report["ok"]
        ok = ok and cyc_ok
        live_bytes = int((((live_s.to(torch.int64) + 15) // 16) * 16).sum().item())
        slots = [alloc.getAvailableSlots(s) for s in probe]
        # free a seeded random half, keep the rest at the front of the list
        mask = torch.from_numpy(rng.random(n) < 0.5).cuda()
        t1 = time.perf_counter()
        alloc.wave_free_batch(offs[:n][mask].contiguous(), tuning=tuning)
        torch.cuda.synchronize()
        t_free = time.perf_counter() - t1
        keep = ~mask
        m = int(keep.sum().item())
        offs[:m] = offs[:n][keep].clone()
        sizes[:m] = sizes[:n][keep].clone()
        count.fill_(m)
        rec = dict(cycle=cyc, new_allocations=n - before, live_allocations=n, fill_fraction=live_bytes / page_bytes,
                   fill_s=t_fill, free_s=t_free, invariants_ok=bool(cyc_ok), slots_after_fill=slots)
        if snap_report is not None:
            rec["snapshot"] = {k: snap_report[k] for k in ("ok", "errors", "used_chunks", "free_pages", "message")}
        results.append(rec)
        if log:
            log(f"oom stress cycle {rec}")
    # free everything that is left; the slot counts come back
    n = int(count.item())
    alloc.wave_free_batch(offs[:n].contiguous(), tuning=tuning)
    torch.cuda.synchronize()
    slots_end = [alloc.getAvailableSlots(s) for s in probe]
    alloc.close()
    return {
        "workload": f"OOM stress: heap {heap_bytes // GiB} GiB, pagesize {pagesize // 1024} KiB, {lanes} lanes fill "
                    f"until {'the first null' if first_null else 'each lane sees a null'} with 80% U[16,256] B / "
                    f"20% U[8,60] KiB, free a random half, {cycles} cycles",
        "cycles": results, "invariants_ok": bool(ok), "slots_restored": slots_end == slots0,
        "mean_fill_fraction": float(np.mean([r["fill_fraction"] for r in results])),
        "tuning": dict(aggregation=tuning.aggregation, class_split=tuning.class_split, regionsize=regionsize,
                       accessblocksize=int(cfg.accessblocksize)),
    }
