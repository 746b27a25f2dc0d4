"""The GPU's serialized mode against the hand-derived known-answer streams.

The streams of oracle_kats.py fix every offset, region fill level and heap
cursor that Scatter.hpp's walk heuristics produce on tiny heaps (derivations
beside each op).  Here the HIP heap runs them through the C ABI
(mmc_serial_stream): each stream whole, checked op by op, and each prefix that
ends at a checkpoint on a fresh heap, whose snapshot must hold the derived
region fill levels, _firstfreeblock and _firstFreePageBased.  No oracle is
involved, so these pin the GPU path on their own.
"""
import numpy as np
import pytest

from oracle_kats import KATS, P, check_geometry

pytestmark = pytest.mark.gpu


def _alloc(mmc, kat, reset):
    g = kat.geometry
    cfg = mmc.ScatterConfig(pagesize=P, regionsize=g.regionsize, accessblocksize=g.accessblocksize,
                            resetfreedpages=reset)
    a = mmc.ScatterAllocator(cfg, g.heap)
    geo = a.geometry()
    check_geometry(g, geo.numpages, geo.numregions, geo.accessblocks, geo.pagesperblock)
    return a


@pytest.mark.parametrize("reset", [False, True], ids=["keep", "resetfreedpages"])
@pytest.mark.parametrize("kat", KATS, ids=[k.name for k in KATS])
def test_gpu_serial_stream_matches_hand_derived(mmc, cuda, kat, reset):
    from mallocmc_amd import snapshot as snap
    torch = cuda
    print(f"kat '{kat.name}' reset={reset}", flush=True)
    ops = np.array(kat.stream(), dtype=np.int64)
    a = _alloc(mmc, kat, reset)
    got = a.serial_stream(torch.from_numpy(ops).cuda()).cpu().numpy().tolist()
    a.close()
    assert got == kat.expected()
    for i, st in sorted(kat.states.items()):
        a = _alloc(mmc, kat, reset)
        a.serial_stream(torch.from_numpy(ops[: i + 1]).cuda())
        s = snap.take(a)
        h = snap.header(s)
        regions = tuple(np.frombuffer(bytes(s), dtype=np.uint32, count=h.geom.numregions,
                                      offset=h.off_regions).tolist())
        assert regions == st.regions, f"after op {i}"
        assert h.firstfreeblock == st.firstfreeblock, f"after op {i}"
        assert h.firstFreePageBased == st.firstFreePageBased, f"after op {i}"
        for size, n in st.slots:
            assert a.getAvailableSlots(size) == n, f"after op {i}"
        a.close()
