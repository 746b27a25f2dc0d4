"""The benchmarked workloads at their full sizes, as GPU tests.

  * mixed-size batch: heap 4 GiB, pagesize 16 KiB, 4096 x 256 = 1,048,576
    lanes, sizes 16 * (1 + splitmix64(seed << 32 | lane) mod 256) B for seeds
    1, 2, 3; every lane's uint64 thread id over its whole buffer, verified;
    frees; nulls counted; free-all restores the slot counts.  In the per-lane
    reference protocol, in the wave path at the same geometry and in the wave
    path at the tuned geometry;
  * many heaps: 64 heaps of 256 MiB, log-uniform 1024..262144 slots, Zipf(1.2)
    sizes 16..65536 B, one captured step replayed 100 times, every slot
    checked after every replay against its seeded coin and size draw, then
    the verify_heap invariants and a snapshot check of every heap;
  * fragmentation / OOM stress: heap 64 GiB, pagesize 64 KiB, 2,097,152
    lanes, 5 fill / free-half cycles, invariants and a snapshot check against
    the complete live list after every fill;
  * serving loop: the whole 30 s Poisson trace on a 32 GiB heap, every
    request's checksum against its closed form, slots restored.
"""
import numpy as np
import pytest

from mallocmc_amd import workloads as wl

pytestmark = pytest.mark.gpu

MiB = 1024 ** 2
GiB = 1024 ** 3
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _no_overlap(torch, offs, sizes):
    live = offs >= 0
    o = offs[live]
    s = ((sizes[live].to(torch.int64) + 15) // 16) * 16
    order = torch.argsort(o)
    o, s = o[order], s[order]
    return o.numel() < 2 or int(((o[:-1] + s[:-1]) > o[1:]).sum().item()) == 0


@pytest.mark.parametrize("form", ["plain", "batched", "batched_tuned"])
def test_mixed_batch_1M_lanes_full_size(mmc, cuda, form):
    torch = cuda
    P, R, xcd = (65536, 16, 8) if form == "batched_tuned" else (16384, 16, 0)
    cfg = mmc.ScatterConfig(pagesize=P, accessblocksize=2 * GiB, regionsize=R, wastefactor=2, resetfreedpages=True)
    a = mmc.ScatterAllocator(cfg, 4 * GiB)
    probe = (16, 496, 4096, P)
    s0 = [a.getAvailableSlots(x) for x in probe]
    n = 4096 * 256
    tun = mmc.Tuning(aggregation=64, xcd_slices=xcd)
    for seed in (1, 2, 3):
        print(f"mixed batch {form} seed {seed}", flush=True)
        sizes = torch.from_numpy(wl.mixed_sizes(seed, n).astype(np.int32)).cuda()
        offs = a.malloc_batch(sizes) if form == "plain" else a.wave_malloc_batch(sizes, tuning=tun)
        a.write_pattern(offs, sizes, salt=0)  # the value written is the lane's id
        r = a.verify_pattern(offs, sizes, salt=0)
        assert r["mismatches"] == 0 and r["misaligned"] == 0 and r["outside"] == 0, r
        assert r["checked"] + r["nulls"] == n
        print(f"  nulls {r['nulls']} of {n}", flush=True)
        if form == "plain":
            # lanes that find no page in their walk come back null: 1.2-3.4% measured
            assert r["nulls"] < n // 20, r
        else:
            assert r["nulls"] == 0, r
        assert _no_overlap(torch, offs, sizes)
        if form == "plain":
            a.free_batch(offs)
        else:
            a.wave_free_batch(offs, tuning=tun)
        torch.cuda.synchronize()
        assert [a.getAvailableSlots(x) for x in probe] == s0
    a.close()


def _mix64(z):
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M64
        return z ^ (z >> np.uint64(31))


def test_multi_heap_full_size_100_replays(mmc, cuda):
    from mallocmc_amd import multiheap, snapshot
    torch = cuda
    work = multiheap.make_workload()
    n = len(work.slot_heap)
    print(f"multi-heap: {len(work.slots_per_heap)} heaps x 256 MiB, {n} slots", flush=True)
    run = multiheap.MultiRun(work, 65536, 32)
    tun = mmc.Tuning(aggregation=64, persistent=False, block_threads=512, xcd_slices=0)
    s0 = [a.getAvailableSlots(64) for a in run.allocs]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            run.multi.step(run.desc, tun, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    slots = np.arange(n, dtype=np.uint64)
    cdf = work.cdf.astype(np.uint64)
    prev_offs = np.full(n, -1, dtype=np.int64)
    prev_sizes = np.zeros(n, dtype=np.int64)
    prev_c = np.zeros(4, dtype=np.int64)
    total_nulls = total_tries = 0
    for step in range(100):
        g.replay()
        torch.cuda.synchronize()
        offs = run.t["offs"].cpu().numpy()
        sizes = run.t["sizes"].cpu().numpy().astype(np.int64)
        c = run.t["counters"].cpu().numpy()
        assert int(run.t["step"].item()) == step + 1
        with np.errstate(over="ignore"):
            r = _mix64(np.uint64(run.desc.seed) ^ _mix64(np.uint64(step) * np.uint64(0x100000001B3) + slots))
        coin = (r % np.uint64(1000)).astype(np.int64)
        k = np.minimum(np.searchsorted(cdf, r >> np.uint64(32), side="left"), len(cdf) - 1)
        want_size = work.table.astype(np.int64)[k]
        prev_live, live = prev_offs >= 0, offs >= 0
        kept = prev_live & (coin >= 500)
        freed = prev_live & (coin < 500)
        tried = ~prev_live & (coin < 500)
        idle = ~prev_live & (coin >= 500)
        assert live[kept].all() and (offs[kept] == prev_offs[kept]).all() and (sizes[kept] == prev_sizes[kept]).all()
        assert not live[freed].any() and not live[idle].any()
        new = tried & live
        assert (sizes[new] == want_size[new]).all(), f"step {step}: size draws"
        dc = c - prev_c
        nulls = int((tried & ~live).sum())
        assert (dc[0], dc[1], dc[2]) == (int(tried.sum()), int(freed.sum()), nulls), f"step {step}: counters"
        total_nulls += nulls
        total_tries += int(tried.sum())
        prev_offs, prev_sizes, prev_c = offs, sizes, c.copy()
    print(f"  nulls {total_nulls} of {total_tries} mallocs", flush=True)
    assert total_nulls < total_tries // 50
    assert run.check_invariants()
    starts = np.concatenate([[0], np.cumsum(work.slots_per_heap)])
    for h, a in enumerate(run.allocs):
        o = run.t["offs"][starts[h]:starts[h + 1]]
        sz = run.t["sizes"][starts[h]:starts[h + 1]]
        rep = snapshot.verify(snapshot.take(a), o, sz, complete=True)
        assert rep["ok"], (h, rep["message"])
        a.wave_free_batch(o.contiguous())
    torch.cuda.synchronize()
    assert [a.getAvailableSlots(64) for a in run.allocs] == s0
    run.close()


def test_oom_stress_full_size(mmc, cuda):
    from mallocmc_amd import oomstress
    r = oomstress.run_default(verify_snapshots=True, log=print)
    for c in r["cycles"]:
        assert c["invariants_ok"], c
        assert c["snapshot"]["ok"], c["snapshot"]
        assert c["slots_after_fill"][3] == 0, "a page was still free at the first null"
        assert 0.6 < c["fill_fraction"] < 0.8, c
    assert r["invariants_ok"] and r["slots_restored"]


def test_serving_full_trace_every_checksum(mmc, cuda):
    from mallocmc_amd import serve
    r = serve.run_default(log=print)
    assert r["requests"] == r["admitted"] == r["completed"] == 60001
    assert r["checksums_checked"] == r["requests"] and r["checksum_mismatches"] == 0
    assert r["nulls"] == 0 and r["slots_restored"]
