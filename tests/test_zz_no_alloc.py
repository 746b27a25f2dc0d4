"""The allocation-free multi-heap step (mmc_multi_step_static) against
mmc_multi_step, through the C ABI.

Both entry points run the same step over the same kind of heaps with the same
descriptor.  Under concurrency, where a request lands depends on wave timing
even for one entry point, so the integer outputs that the seeded coins fix
are compared bit-exactly -- which slots hold an allocation after every step,
each live slot's size, the malloc/free/null/tile counters and the step
counter -- and the placements are held to the verify_heap invariants (no
overlap, aligned, inside the heap, patterns intact) and the snapshot checker.
Two calls in a row, the second with fewer heaps (and different heaps, so a
stale table entry from the first call would land requests in the wrong
heap), then graph replays of a captured static step.
"""
import numpy as np
import pytest

from mallocmc_amd import snapshot

pytestmark = pytest.mark.gpu

MiB = 1024 ** 2


def _setup(torch, mmc, nheaps, per, seed, heap_bytes=64 * MiB, pagesize=16384):
    from test_gpu_wave import _step_setup
    allocs, multi, dev, d, slot_heap = _step_setup(torch, mmc, nheaps, per, heap_bytes, pagesize, seed=seed)
    return allocs, multi, dev, d, slot_heap


def _state(dev):
    offs = dev["offs"].cpu().numpy()
    return offs >= 0, dev["sizes"].cpu().numpy(), dev["counters"].cpu().numpy().copy(), int(dev["step"].item())


def _check_heaps(torch, allocs, dev, slot_heap):
    from test_gpu_wave import no_overlap
    offs, sizes = dev["offs"], dev["sizes"]
    for hidx, a in enumerate(allocs):
        sel = torch.from_numpy(np.nonzero(slot_heap == hidx)[0]).cuda()
        o, s = offs[sel].contiguous(), sizes[sel].contiguous()
        assert no_overlap(torch, o, s), f"heap {hidx}"
        a.write_pattern(o, s, salt=hidx + 1)
        r = a.verify_pattern(o, s, salt=hidx + 1)
        assert r["mismatches"] == 0 and r["misaligned"] == 0 and r["outside"] == 0, (hidx, r)
        padded = ((s.to(torch.int64) + 15) // 16) * 16
        rep = snapshot.verify(snapshot.take(a), o, padded, complete=True)
        assert rep["ok"], (hidx, rep["message"])


def _free_all(torch, allocs, dev, slot_heap):
    for hidx, a in enumerate(allocs):
        sel = torch.from_numpy(np.nonzero(slot_heap == hidx)[0]).cuda()
        a.wave_free_batch(dev["offs"][sel].contiguous())
    torch.cuda.synchronize()


def _compare(torch, mmc, nheaps, per, seed, steps, tun):
    print(f"no-alloc step: {nheaps} heaps x {per} slots, seed {seed}, persistent={tun.persistent}", flush=True)
    ref_allocs, ref_multi, ref_dev, ref_d, slot_heap = _setup(torch, mmc, nheaps, per, seed)
    new_allocs, new_multi, new_dev, new_d, _ = _setup(torch, mmc, nheaps, per, seed)
    new_multi.close()  # the static step needs no group
    s0 = [a.getAvailableSlots(64) for a in new_allocs]
    for _ in range(steps):
        ref_multi.step(ref_d, tun)
        mmc.multi_step_static(new_allocs, new_d, tun)
        ra, rs, rc, rstep = _state(ref_dev)
        na, ns, nc, nstep = _state(new_dev)
        np.testing.assert_array_equal(na, ra)
        np.testing.assert_array_equal(ns[na], rs[ra])
        np.testing.assert_array_equal(nc, rc)
        assert nstep == rstep
    assert rc[2] == 0 and rc[0] > 0 and rc[1] > 0
    assert int(new_dev["cursor"].abs().sum().item()) == 0, "the caller's cursor must stay untouched"
    _check_heaps(torch, new_allocs, new_dev, slot_heap)
    _free_all(torch, new_allocs, new_dev, slot_heap)
    assert [a.getAvailableSlots(64) for a in new_allocs] == s0
    ref_multi.close()
    for a in ref_allocs + new_allocs:
        a.close()


@pytest.mark.parametrize("persistent,block_threads", [(True, 0), (False, 512)])
def test_static_step_matches_multi_step_twice(mmc, cuda, persistent, block_threads):
    torch = cuda
    tun = mmc.Tuning(persistent=persistent, aggregation=64, block_threads=block_threads)
    _compare(torch, mmc, 8, 4096, 21, 12, tun)
    _compare(torch, mmc, 3, 1500, 22, 12, tun)  # fewer heaps, ragged tiles, other heaps


def test_static_step_rejects_too_many_heaps(mmc, cuda):
    torch = cuda
    allocs, multi, dev, d, _ = _setup(torch, mmc, 1, 64, 3, heap_bytes=4 * MiB, pagesize=4096)
    with pytest.raises(mmc.MallocMCError, match="1 to 256 heaps"):
        mmc.multi_step_static(allocs * (mmc.STATIC_MAX_HEAPS + 1), d)
    multi.close()
    allocs[0].close()


def test_static_step_graph_replays(mmc, cuda):
    """A captured static step replays like eager mmc_multi_step calls."""
    torch = cuda
    tun = mmc.Tuning(persistent=False, aggregation=64, block_threads=512)
    ref_allocs, ref_multi, ref_dev, ref_d, slot_heap = _setup(torch, mmc, 4, 4096, 31)
    new_allocs, new_multi, new_dev, new_d, _ = _setup(torch, mmc, 4, 4096, 31)
    new_multi.close()
    mmc.multi_step_static(new_allocs, new_d, tun)  # first call on the device, outside the capture
    ref_multi.step(ref_d, tun)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            mmc.multi_step_static(new_allocs, new_d, tun, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for i in range(8):
        g.replay()
        ref_multi.step(ref_d, tun)
        torch.cuda.synchronize()
        _ = new_dev["offs"].sum().item()  # eager work between replays
        ra, rs, rc, rstep = _state(ref_dev)
        na, ns, nc, nstep = _state(new_dev)
        np.testing.assert_array_equal(na, ra, err_msg=f"replay {i}")
        np.testing.assert_array_equal(ns[na], rs[ra], err_msg=f"replay {i}")
        np.testing.assert_array_equal(nc, rc, err_msg=f"replay {i}")
        assert nstep == rstep
    _check_heaps(torch, new_allocs, new_dev, slot_heap)
    ref_multi.close()
    for a in ref_allocs + new_allocs:
        a.close()
