"""Host logic of the multi-heap driver (mallocmc_amd/multiheap.py), without a
GPU: the workload's tiles, and the two-stage autotune's choice over a stand-in
for the device run whose rate is a known function of the tuning."""
import numpy as np
import pytest

from mallocmc_amd import multiheap as mh


@pytest.mark.parametrize("small_first", [True, False])
@pytest.mark.parametrize("pad", [1, 4, 16])
def test_workload_tiles_cover_each_heap_once(pad, small_first):
    work = mh.make_workload(n_heaps=8, lo=100, hi=1000, pad=pad, small_first=small_first)
    starts = np.concatenate([[0], np.cumsum(work.slots_per_heap)])
    # every slot of a heap lies in exactly one tile of that heap, tiles start 64 apart
    covered = np.zeros(len(work.slot_heap), dtype=np.int64)
    for t in work.tiles:
        if t == mh.NO_TILE:
            continue
        h = work.slot_heap[t]
        end = min(int(t) + 64, int(starts[h + 1]))
        covered[int(t):end] += 1
        assert (work.slot_heap[int(t):end] == h).all()
    assert (covered == 1).all()
    # padded to `pad` entries per heap: every aligned group of `pad` entries
    # (a workgroup of pad, or a divisor of pad, waves) lies in one heap
    assert len(work.tiles) % pad == 0
    for g in range(0, len(work.tiles), pad):
        hs = {int(work.slot_heap[t]) for t in work.tiles[g:g + pad] if t != mh.NO_TILE}
        assert len(hs) <= 1, (g, hs)
    assert work.cdf[-1] == 2 ** 32 - 1 and (np.diff(work.cdf.astype(np.int64)) >= 0).all()
    # heaps appear in one block each, smallest slot count first (else in heap order)
    heaps = [int(work.slot_heap[t]) for t in work.tiles if t != mh.NO_TILE]
    blocks = [h for i, h in enumerate(heaps) if i == 0 or heaps[i - 1] != h]
    assert sorted(blocks) == list(range(8))
    if small_first:
        assert [work.slots_per_heap[h] for h in blocks] == sorted(work.slots_per_heap)
    else:
        assert blocks == list(range(8))


class _FakeRun:
    """Stands in for MultiRun: a step serves 1000 mallocs and 1000 frees in a
    time that depends on the geometry and the tuning, with one optimum."""
    checked = 0

    def __init__(self, work, pagesize, regionsize, heap_bytes=None, seed=7):
        self.P, self.R = pagesize, regionsize

    def run(self, steps, tun, graph=True):
        ms = 1.0
        ms += 0.0 if self.P == 65536 else 0.5
        ms += 0.0 if self.R == 32 else 0.25
        ms += 0.0 if not tun.persistent else 0.4
        ms += {0: 0.0, 256: 0.1, 512: 0.0, 1024: 0.2}[tun.block_threads]
        ms += 0.0 if tun.aggregation == 16 else 0.05
        ms += 0.0 if tun.xcd_slices == 0 else 0.02
        ms += 0.0 if not tun.fresh_single else 0.03
        return ms * steps, np.array([1000 * steps, 1000 * steps, 0, 10 * steps])

    def check_invariants(self):
        _FakeRun.checked += 1
        return True

    def close(self):
        pass


def test_two_stage_autotune_finds_the_optimum(monkeypatch):
    monkeypatch.setattr(mh, "MultiRun", _FakeRun)
    out = mh.run_default(steps=4, tune_steps=2)
    best = out["tuned"]
    assert (best["pagesize"], best["regionsize"]) == (65536, 32)
    assert (best["persistent"], best["block_threads"]) == (False, 512)
    # the second stage at the best point finds the aggregation, XCD slicing
    # and multi-page claims of unformatted pages
    assert (best["aggregation"], best["xcd_slices"], best["fresh_single"]) == (16, 0, False)
    # 3 pagesizes x 3 regionsizes x 4 launch forms, then 3 x 2 x 2 - 1 second-stage points
    assert out["autotune_points"] == 9 * len(mh.LAUNCH_FORMS) + 11
    assert out["invariants_ok"] and _FakeRun.checked >= 1
    assert np.isclose(out["mallocs_frees_per_s"], 2000 / (1.0 / 1e3))
