"""GPU test of the fragmentation / OOM stress cycle (config 4 shape on a
1 GiB heap): fill until null, free a random half, refill; invariants after
every fill and the slot counts restored at the end."""
import pytest

pytestmark = pytest.mark.gpu

GiB = 1024 ** 3


@pytest.mark.parametrize("split,first_null", [(0, False), (2, False), (2, True)])
def test_oom_stress_cycles(mmc, cuda, split, first_null):
    from mallocmc_amd import oomstress
    r = oomstress.run_config4(heap_bytes=1 * GiB, lanes=65536, cycles=3,
                              tuning=mmc.Tuning(aggregation=64, class_split=split), log=print,
                              first_null=first_null)
    assert r["invariants_ok"], r
    assert r["slots_restored"], r
    for c in r["cycles"]:
        assert c["fill_fraction"] > 0.5, c
