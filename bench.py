#!/usr/bin/env python3
"""Headline benchmark: mallocs+frees/s at 1,048,576 lanes, batched vs plain, one MI355X.

Workload (the mixed-size batch): heap 4 GiB, 1,048,576 requests per
step with sizes 16 * (1 + splitmix64(seed << 32 | lane) mod 256) B (uniform
16..4096 B), seeds 1, 2, 3 in turn; a step is one malloc launch over all
requests followed by one free launch over the results.  The size arrays sit
in HBM before timing starts.

Forms:
  * plain    -- the reference's per-lane ScatterAlloc protocol, one lane per
                request, one launch per phase, pagesize 16 KiB, regionsize 16;
  * batched  -- wave-cooperative claims with the geometry and wave tuning the
                autotuner picks, malloc+free of a step captured in one hipGraph
                and replayed.  This is `value`.  The same form on the plain
                form's geometry (default tuning, eager launches) is reported
                beside it as batched_reference_geometry.

Also reported: the device-atomic ceiling (microbenchmark), the roofline of
the dominant kernel, the CPU baseline (the oracle on this host's cores, the
reference's CpuOmp2Blocks execution model with 16-way mutex atomics), the
serving-loop p50/p99 step latency (mallocmc_amd/serve.py), the
multi-heap batched form (mallocmc_amd/multiheap.py) and the
fragmentation/OOM stress cycles (mallocmc_amd/oomstress.py).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "mallocs+frees/s at 1M threads, batched vs plain, 1 GPU; p50/p99 serve step us"
GiB = 1024 ** 3
N_LANES = 1 << 20
HEAP = 4 * GiB
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one malloc request: 4 B size in, 8 B offset out, and the
# reference protocol's three 4-byte metadata read-modify-writes (chunk size CAS,
# count add, bitmask or) at 8 B each
MALLOC_BYTES_PER_REQUEST = 4 + 8 + 3 * 8
MALLOC_RMW_PER_REQUEST = 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def atomic_ceiling(torch, mmc):
    """Agent-scope returning atomicAdd, one 128-byte line per lane, full chip."""
    L = mmc.lib()
    words = 1 << 26
    buf = torch.zeros(words, dtype=torch.int32, device="cuda")
    sink = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    blocks, threads, iters = 8192, 256, 16
    best = 0.0
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = L.mmc_bench_atomics(ctypes.c_void_p(buf.data_ptr()), words, 0, iters, blocks, threads,
                                 ctypes.c_void_p(sink.data_ptr()), s)
        e1.record()
        torch.cuda.synchronize()
        if rc != 0:
            raise RuntimeError("atomics microbenchmark failed")
        best = max(best, blocks * threads * iters / (e0.elapsed_time(e1) / 1e3))
    del buf
    return best


def measure_plain(torch, mmc, batches, steps):
    cfg = mmc.ScatterConfig(pagesize=16384, accessblocksize=2 * GiB, regionsize=16, resetfreedpages=True)
    a = mmc.ScatterAllocator(cfg, HEAP)
    offs = torch.empty(N_LANES, dtype=torch.int64, device="cuda")
    # one untimed step
    a.malloc_batch(batches[0], offs)
    a.free_batch(offs)
    torch.cuda.synchronize()
    total_ms, nulls, mall_ms = 0.0, 0, 0.0
    for i in range(steps):
        sizes = batches[i % len(batches)]
        e0, e1, e2 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e0.record()
        a.malloc_batch(sizes, offs)
        e1.record()
        a.free_batch(offs)
        e2.record()
        torch.cuda.synchronize()
        nulls += int((offs < 0).sum().item())
        total_ms += e0.elapsed_time(e2)
        mall_ms += e0.elapsed_time(e1)
    a.close()
    return dict(value=2 * N_LANES * steps / (total_ms / 1e3), ms_per_step=total_ms / steps,
                malloc_ms=mall_ms / steps, nulls=nulls, steps=steps,
                config={"pagesize": 16384, "regionsize": 16, "path": "per-lane", "launch": "plain"})


def measure_batched_reference_geometry(torch, mmc, batches, steps):
    """The batched form on the plain form's configuration (pagesize 16 KiB,
    regionsize 16, the reference's values for this workload, no pagesize-leak
    fix), default tuning: what the wave path gives without the autotuner."""
    cfg = mmc.ScatterConfig(pagesize=16384, accessblocksize=2 * GiB, regionsize=16, resetfreedpages=True)
    a = mmc.ScatterAllocator(cfg, HEAP)
    tun = mmc.Tuning()
    offs = torch.empty(N_LANES, dtype=torch.int64, device="cuda")
    for b in batches:
        a.wave_malloc_batch(b, offs, tuning=tun)
        a.wave_free_batch(offs, tuning=tun)
    torch.cuda.synchronize()
    total_ms, nulls = 0.0, 0
    for i in range(steps):
        sizes = batches[i % len(batches)]
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        a.wave_malloc_batch(sizes, offs, tuning=tun)
        a.wave_free_batch(offs, tuning=tun)
        e1.record()
        torch.cuda.synchronize()
        total_ms += e0.elapsed_time(e1)
        if i < len(batches):
            nulls += int((offs < 0).sum().item())
    a.close()
    return dict(value=2 * N_LANES * steps / (total_ms / 1e3), ms_per_step=total_ms / steps, steps=steps,
                nulls_first_steps=nulls,
                config={"pagesize": 16384, "regionsize": 16, "aggregation": 64, "xcd_slices": 0,
                        "path": "wave-cooperative", "launch": "eager, one malloc + one free launch per step"})


def cpu_baseline(sizes_np):
    """The oracle (reference CpuOmp2Blocks model, 16-way mutex atomics) on this host."""
    from oracle import oracle as orc
    orc.lib()
    n = len(sizes_np)
    cores = min(16, os.cpu_count() or 1)
    cfg = orc.make_config(pagesize=16384, accessblocksize=2 * GiB, resetfreedpages=True)
    h = orc.OracleHeap(cfg, HEAP)
    h.set_atomic_mode(orc.ATOMIC_MUTEX16)
    res, _ = h.bench_concurrent(sizes_np, nthreads=cores)
    h.close()
    ops = 2 * n / (res.t_malloc + res.t_free)
    return {"value": ops, "unit": "mallocs+frees/s", "cores": res.threads_used, "kind": "port",
            "sample": f"mixed-size batch, seed 1: {n} requests, 4 GiB heap, pagesize 16 KiB, one malloc and one free "
                      f"each, blocks x 1 thread over OpenMP (CpuOmp2Blocks), 16-way mutex atomics",
            "nulls": int(res.nulls), "mismatches": int(res.mismatches),
            "t_malloc_s": res.t_malloc, "t_free_s": res.t_free}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="use the stored best point instead of tuning")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-multi", action="store_true", help="skip the multi-heap section")
    ap.add_argument("--no-oom", action="store_true", help="skip the OOM-stress section")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the heap is per-device state and this build runs on one MI355X; "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    import torch
    import mallocmc_amd as mmc
    from mallocmc_amd import autotune
    from mallocmc_amd import workloads as wl

    mmc.lib()
    torch.cuda.init()
    seeds = (1, 2, 3)
    sizes_np = [wl.mixed_sizes(s, N_LANES) for s in seeds]
    batches = [torch.from_numpy(x.astype(np.int32)).cuda() for x in sizes_np]
    torch.cuda.synchronize()

    ceiling = atomic_ceiling(torch, mmc)
    log(f"atomic ceiling {ceiling:.3e}/s")

    # the batched form runs with the native fix for pagesize-sized chunks (their
    # pages are released through the chunked path instead of leaking a slot)
    base = mmc.ScatterConfig(accessblocksize=2 * GiB, resetfreedpages=True, flags=mmc.FLAG_FIX_PAGESIZE_LEAK)
    if args.no_tune:
        # the point the autotuner picks on this workload (pagesize 64 KiB, regionsize 32)
        best = autotune.TunePoint(pagesize=65536, regionsize=32, aggregation=64, persistent=False, xcd_slices=0)
        tuned_points = 0
    else:
        t0 = time.time()
        # persistent launches never win on this one-heap workload: they serve
        # requests in arrival order, without the block partition (three to
        # four times slower); the batched form sweeps the rest
        space = dict(autotune.DEFAULT_SPACE, persistent=(False,))
        best, points = autotune.tune(base, HEAP, batches, space=space, log=log)
        tuned_points = len(points)
        log(f"autotune: {len(points)} points in {time.time() - t0:.1f}s, best {best}")
    cfg = best.config(base)
    tun = best.tuning()

    alloc = mmc.ScatterAllocator(cfg, HEAP)
    slots0 = alloc.getAvailableSlots(16)
    offs = torch.empty(N_LANES, dtype=torch.int64, device="cuda")
    stream = torch.cuda.Stream()
    graphs = []
    with torch.cuda.stream(stream):
        # eager warm-up of every batch before capture
        for b in batches:
            alloc.wave_malloc_batch(b, offs, tuning=tun, stream=stream)
            alloc.wave_free_batch(offs, tuning=tun, stream=stream)
        stream.synchronize()
        for b in batches:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=stream):
                alloc.wave_malloc_batch(b, offs, tuning=tun, stream=stream)
                alloc.wave_free_batch(offs, tuning=tun, stream=stream)
            graphs.append(g)
    torch.cuda.synchronize()
    for i in range(args.warmup):
        graphs[i % 3].replay()
    torch.cuda.synchronize()

    # timed region: K steps, barrier + synchronize on both sides
    # graph replays launch on the current stream: the events go there too
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        graphs[i % 3].replay()
    ev1.record()
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    gpu_ms = ev0.elapsed_time(ev1)
    value = 2 * N_LANES * args.steps / (gpu_ms / 1e3)
    log(f"batched: {args.steps} steps {gpu_ms:.3f} ms (wall {t_wall * 1e3:.3f} ms) -> {value:.3e}/s")

    # dominant kernel: the wave malloc launch, timed alone on its stream
    mal_ms = []
    nulls = 0
    for i in range(6):
        b = batches[i % 3]
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        alloc.wave_malloc_batch(b, offs, tuning=tun)
        e1.record()
        torch.cuda.synchronize()
        mal_ms.append(e0.elapsed_time(e1))
        if i == 0:
            alloc.write_pattern(offs, b, salt=3)
            v = alloc.verify_pattern(offs, b, salt=3)
            nulls = v["nulls"]
            invariants_ok = v["mismatches"] == 0 and v["misaligned"] == 0 and v["outside"] == 0
        alloc.wave_free_batch(offs, tuning=tun)
    torch.cuda.synchronize()
    slots_restored = alloc.getAvailableSlots(16) == slots0
    alloc.close()
    del graphs
    mal_avg_ms = float(np.mean(mal_ms[1:]))
    # HBM bytes per malloc launch from the committed counter passes (see the file)
    traffic = None
    tpath = os.path.join(HERE, "profiles", "batched_mixed1M_hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("waveMallocKernel<NoopDistribution>", {}).get("hbm_bytes_per_launch")
    achieved_gbs = N_LANES * MALLOC_BYTES_PER_REQUEST / (mal_avg_ms / 1e3) / 1e9
    atom_rate = N_LANES * MALLOC_RMW_PER_REQUEST / (mal_avg_ms / 1e3)
    # the reference protocol's read-modify-writes per malloc+free pair, uncontended,
    # with resetfreedpages: malloc = chunk-size CAS + count add + mask or + chunk-size
    # re-CAS (Scatter.hpp:469-517, 589-591); free = mask and (+ segment-hint and on
    # a hierarchical page) + failed count CAS + count sub (Scatter.hpp:646-716):
    # 8 for chunks up to the hierarchy threshold, 7 above it
    ht = (best.pagesize - 8) // 33
    sz = np.concatenate([b.cpu().numpy() for b in batches])
    frac_h = float(np.mean(((sz + 15) // 16 * 16) <= ht))
    rmw_pair = 8.0 * frac_h + 7.0 * (1.0 - frac_h)
    pair_rate = value / 2.0 * rmw_pair

    # copy kernel of known size: calibrates the HBM byte counters of a profile run
    copy_bytes = 256 * 1024 * 1024
    src = torch.empty(copy_bytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    if mmc.lib().mmc_bench_copy(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), copy_bytes,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)) != 0:
        raise RuntimeError("copy kernel failed")
    torch.cuda.synchronize()
    del src, dst

    ref_geom = measure_batched_reference_geometry(torch, mmc, batches, args.steps)
    log(f"batched at the reference geometry: {ref_geom}")

    plain = None
    if not args.no_plain:
        plain = measure_plain(torch, mmc, batches, args.plain_steps)
        log(f"plain: {plain}")

    multi = None
    if not args.no_multi:
        from mallocmc_amd import multiheap
        multi = multiheap.run_default(log=log, tune_steps=10)

    oom = None
    if not args.no_oom:
        from mallocmc_amd import oomstress
        oom = oomstress.run_default(log=log)

    serve = None
    if not args.no_serve:
        try:
            from mallocmc_amd import serve as serve_mod
            serve = serve_mod.run_default(log=log)
        except ImportError:
            serve = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sizes_np[0])
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "mallocs+frees/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu_ms / args.steps,
        "wall_ms_per_step": t_wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "mixed-size batch: 1,048,576 requests/step, sizes 16..4096 B (splitmix64 seeds 1,2,3), 4 GiB heap",
            "form": "batched+autotuned (wave-cooperative, hipGraph malloc+free per step)",
            "pagesize": best.pagesize, "regionsize": best.regionsize, "aggregation": best.aggregation,
            "persistent": best.persistent, "xcd_slices": best.xcd_slices, "resetfreedpages": True,
            "fix_pagesize_leak": True,
            "autotune_points": tuned_points,
        },
        "nulls_per_step": nulls,
        "invariants_ok": bool(invariants_ok),
        "slots_restored": bool(slots_restored),
        "batched_reference_geometry": ref_geom,
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "roofline": {
            "bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "waveMallocKernel", "avg_launch_ms": mal_avg_ms,
            "bytes_per_request": MALLOC_BYTES_PER_REQUEST,
        },
        "atomic_roofline": {
            "achieved": atom_rate, "peak": ceiling, "unit": "reference-protocol RMW/s",
            "frac": atom_rate / ceiling, "rmw_per_request": MALLOC_RMW_PER_REQUEST,
            "per_pair": {
                "achieved": pair_rate, "frac": pair_rate / ceiling, "rmw_per_pair": rmw_pair,
                "hierarchical_fraction": frac_h,
                "note": "whole step (malloc+free pairs/s) x the reference protocol's RMW per pair; "
                        "the wave path issues fewer RMW than that, so frac may exceed 1",
            },
        },
        "cpu_baseline": cpu,
        "serve": serve,
        "batched_multi_heap": multi,
        "oom_stress": oom,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
