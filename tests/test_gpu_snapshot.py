"""GPU tests of heap snapshots (include/mallocmc_snapshot.h): the state the
wave path leaves on the device passes the host checker with the complete
live list, at every stage of a malloc / free-half / free-all cycle, and the
snapshot round-trips through a file and carries the payload."""
import os

import numpy as np
import pytest

from mallocmc_amd import workloads as wl

pytestmark = pytest.mark.gpu

MiB = 1024 ** 2


def _sizes(torch, n, seed, P):
    """Mixed sizes with a few multi-page requests."""
    s = wl.mixed_sizes(seed, n).astype(np.int64)
    rng = np.random.default_rng(seed)
    big = rng.random(n) < 0.002
    s[big] = rng.integers(P + 16, 3 * P, size=int(big.sum()))
    return torch.from_numpy(s.astype(np.int32)).cuda()


def test_snapshot_verifies_through_a_cycle(mmc, cuda, tmp_path):
    from mallocmc_amd import snapshot as S
    torch = cuda
    P = 16384
    cfg = mmc.ScatterConfig(pagesize=P, accessblocksize=512 * MiB, regionsize=16, resetfreedpages=True,
                            flags=mmc.FLAG_FIX_PAGESIZE_LEAK)
    a = mmc.ScatterAllocator(cfg, 512 * MiB)
    n = 65536
    sizes = _sizes(torch, n, 5, P)
    print(f"snapshot cycle n={n}", flush=True)
    offs = a.wave_malloc_batch(sizes, tuning=mmc.Tuning(aggregation=64))
    torch.cuda.synchronize()
    padded = ((sizes.to(torch.int64) + 15) // 16 * 16).to(torch.int32)
    snap = S.take(a)
    r = S.verify(snap, offs, padded, complete=True)
    assert r["ok"], r
    assert r["live_checked"] == int((offs >= 0).sum()) and r["runs"] > 0
    # free a seeded half with the wave path, the rest stays live
    keep = torch.from_numpy(np.random.default_rng(1).random(n) < 0.5).cuda()
    a.wave_free_batch(offs[~keep].contiguous())
    torch.cuda.synchronize()
    snap = S.take(a)
    r = S.verify(snap, offs[keep], padded[keep], complete=True)
    assert r["ok"], r
    # save / load round trip
    path = os.path.join(tmp_path, "heap.mmcsnap")
    S.save(path, snap)
    assert S.load(path) == snap
    # free everything: no chunked page, no run, nothing formatted is left
    a.wave_free_batch(offs[keep].contiguous())
    torch.cuda.synchronize()
    r = S.verify(S.take(a))
    assert r["ok"] and r["chunked_pages"] == 0 and r["run_pages"] == 0 and r["used_chunks"] == 0, r
    a.close()


def test_snapshot_payload_holds_the_pattern(mmc, cuda):
    from mallocmc_amd import snapshot as S
    torch = cuda
    cfg = mmc.ScatterConfig(pagesize=4096, accessblocksize=64 * MiB, regionsize=16, resetfreedpages=True)
    a = mmc.ScatterAllocator(cfg, 64 * MiB)
    n = 4096
    sizes = torch.from_numpy(wl.mixed_sizes(3, n, steps=16).astype(np.int32)).cuda()
    offs = a.wave_malloc_batch(sizes)
    a.write_pattern(offs, sizes, salt=1000)
    torch.cuda.synchronize()
    snap = S.take(a, payload=True)
    hdr = S.header(snap)
    assert hdr.off_payload > 0
    o = offs.cpu().numpy()
    z = sizes.cpu().numpy()
    for i in (0, n // 2, n - 1):
        words = np.frombuffer(bytes(snap[hdr.off_payload + int(o[i]): hdr.off_payload + int(o[i]) + int(z[i]) // 8 * 8]),
                              dtype=np.uint64)
        assert (words == 1000 + i).all()
    assert S.verify(snap, offs, sizes, complete=True)["ok"]
    a.close()


@pytest.mark.parametrize("persistent,block_threads,fresh_single", [(True, 0, False), (False, 512, False),
                                                                   (False, 512, True)])
def test_snapshot_after_multi_heap_steps(mmc, cuda, persistent, block_threads, fresh_single):
    """Concurrent mallocs and frees of the multi-heap step (persistent, or one
    partitioning workgroup per 8 tiles; unformatted pages claimed several per
    round trip or one at a time) leave every heap consistent with the slots'
    live allocations."""
    from mallocmc_amd import snapshot as S
    from test_gpu_wave import _step_setup
    torch = cuda
    nheaps, per = 4, 8192
    allocs, multi, dev, d, slot_heap = _step_setup(torch, mmc, nheaps, per, 128 * MiB, 4096, seed=13, nsizes=13)
    tun = mmc.Tuning(persistent=persistent, aggregation=64, block_threads=block_threads, fresh_single=fresh_single)
    print(f"snapshot after multi-heap steps persistent={persistent} block_threads={block_threads} "
          f"fresh_single={fresh_single}", flush=True)
    for _ in range(10):
        multi.step(d, tun)
    torch.cuda.synchronize()
    offs, sizes = dev["offs"], dev["sizes"]
    for hidx, a in enumerate(allocs):
        sel = torch.from_numpy(np.nonzero(slot_heap == hidx)[0]).cuda()
        o, s = offs[sel], sizes[sel]
        live = o >= 0
        r = S.verify(S.take(a), o[live], ((s[live].to(torch.int64) + 15) // 16 * 16), complete=True)
        assert r["ok"], (hidx, r)
