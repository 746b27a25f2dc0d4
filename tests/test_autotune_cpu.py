"""The one-heap autotuner's choice logic on the CPU, over a stand-in device:
the leaders of the sweep are measured again before the choice, and a
preferred point within the tie margin of the fastest wins."""
import mallocmc_amd as mmc
from mallocmc_amd import autotune

# true rates of the stand-in device; the sweep's single pass sees them with
# noise that makes (16384, 16) look fastest
TRUE = {(65536, 16, 8): 1.00e10, (65536, 32, 0): 1.02e10, (65536, 16, 0): 0.99e10, (16384, 16, 8): 0.90e10}
SWEEP_NOISE = {(16384, 16, 8): 1.3}


class _FakeAlloc:
    def __init__(self, cfg, heap_bytes):
        self.cfg = cfg

    def getAvailableSlots(self, size):
        return 100

    def close(self):
        pass


def _fake_measure(alloc, batches, tuning, rounds=2, timed=1):
    key = (alloc.cfg.pagesize, alloc.cfg.regionsize, tuning.xcd_slices)
    base = TRUE.get(key, 1e9) if tuning.aggregation == 64 else 1e8
    noise = SWEEP_NOISE.get(key, 1.0) if timed == 1 else 1.0
    return base * noise, 0, True


def _space():
    return dict(pagesize=(16384, 65536), regionsize=(16, 32), aggregation=(16, 64), persistent=(False,),
                xcd_slices=(0, 8))


def test_refine_removes_a_noisy_sweep_winner(monkeypatch):
    monkeypatch.setattr(mmc, "ScatterAllocator", _FakeAlloc)
    monkeypatch.setattr(autotune, "measure", _fake_measure)
    base = mmc.ScatterConfig(resetfreedpages=True)
    best, points = autotune.tune(base, 1 << 30, [], space=_space())
    assert (best.pagesize, best.regionsize, best.xcd_slices) == (65536, 32, 0)
    assert len(points) == 2 * 2 * 2 * 2


def test_preferred_point_wins_a_tie(monkeypatch):
    monkeypatch.setattr(mmc, "ScatterAllocator", _FakeAlloc)
    monkeypatch.setattr(autotune, "measure", _fake_measure)
    base = mmc.ScatterConfig(resetfreedpages=True)
    prefer = autotune.TunePoint(pagesize=65536, regionsize=16, aggregation=64, persistent=False, xcd_slices=8)
    best, _ = autotune.tune(base, 1 << 30, [], space=_space(), prefer=prefer, tie=0.03)
    assert (best.pagesize, best.regionsize, best.xcd_slices) == (65536, 16, 8)
    # outside the margin the fastest point is kept
    best, _ = autotune.tune(base, 1 << 30, [], space=_space(), prefer=prefer, tie=0.01)
    assert (best.pagesize, best.regionsize, best.xcd_slices) == (65536, 32, 0)


def test_requests_per_lane_are_swept_for_non_persistent_points_only(monkeypatch):
    monkeypatch.setattr(mmc, "ScatterAllocator", _FakeAlloc)
    seen = []

    def measure(alloc, batches, tuning, rounds=2, timed=1):
        seen.append((tuning.persistent, tuning.lane_requests))
        return (2e9 if tuning.lane_requests == 2 else 1e9), 0, True

    monkeypatch.setattr(autotune, "measure", measure)
    base = mmc.ScatterConfig(resetfreedpages=True)
    space = dict(pagesize=(65536,), regionsize=(16,), aggregation=(64,), persistent=(False, True),
                 xcd_slices=(0,), lane_requests=(1, 2, 4))
    best, points = autotune.tune(base, 1 << 30, [], space=space)
    # persistent launches serve one request per lane: one point, not three
    assert len(points) == 3 + 1
    assert {(p.persistent, p.lane_requests) for p in points} == {(False, 1), (False, 2), (False, 4), (True, 1)}
    assert best.lane_requests == 2 and not best.persistent
    assert best.tuning().lane_requests == 2
