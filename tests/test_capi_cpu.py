"""CPU-only checks of the C ABI library: it loads, exports every entry point
the headers declare, and its host geometry equals the oracle's."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", h) for h in sorted(os.listdir(os.path.join(REPO, "include")))
           if h.endswith(".h")]

MiB = 1024 ** 2
GiB = 1024 ** 3


def declared_symbols():
    syms = set()
    for h in HEADERS:
        text = open(h).read()
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(mmc_[a-z_0-9]+)\s*\(", text, re.M):
            syms.add(m.group(1))
    return sorted(syms)


def test_headers_declare_entry_points():
    syms = declared_symbols()
    assert "mmc_create" in syms and "mmc_malloc_batch" in syms and "mmc_serial_stream" in syms
    assert len(syms) >= 20


def test_library_exports_every_declared_symbol(mmc):
    L = mmc.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_lists_all_symbols(mmc):
    assert set(declared_symbols()) <= set(mmc.EXPORTED_SYMBOLS) | set(getattr(mmc, "BATCHED_SYMBOLS", []))


def test_default_config_matches_reference_defaults(mmc):
    c = mmc.MmcConfig()
    assert mmc.lib().mmc_default_config(ctypes.byref(c)) == 0
    # Scatter.hpp:56-85, Shrink.hpp:60-66
    assert (c.pagesize, c.accessblocksize, c.regionsize, c.wastefactor, c.resetfreedpages) == (
        4096, 2 * GiB, 16, 2, 0)
    assert (c.hashingK, c.hashingDistMP, c.hashingDistWP, c.hashingDistWPRel) == (38183, 17497, 1, 1)
    assert c.dataAlignment == 16


@pytest.mark.parametrize("pagesize,ab,region,heap", [
    (4096, 2 * MiB, 16, 64 * MiB),
    (16384, 2 * GiB, 16, 4 * GiB),
    (16384, 2 * MiB, 8, 256 * MiB),
    (65536, 64 * MiB, 32, 256 * MiB),
    (65536, 2 * GiB, 16, 64 * GiB),
    (4096, 2 * GiB, 16, 32 * GiB),
    (4096, 2 * GiB, 16, 1 * GiB),     # clamped (D1)
    (1000, 100000, 3, 5 * MiB + 17),  # odd sizes
])
def test_host_geometry_equals_oracle(mmc, oracle_mod, pagesize, ab, region, heap):
    cfg = mmc.ScatterConfig(pagesize=pagesize, accessblocksize=ab, regionsize=region)
    g = mmc.compute_geometry(cfg, heap)
    og = oracle_mod.geometry(oracle_mod.make_config(pagesize=pagesize, accessblocksize=ab, regionsize=region), heap)
    for f, _ in mmc.MmcGeometry._fields_:
        assert getattr(g, f) == getattr(og, f), f


def test_invalid_config_is_rejected(mmc):
    cfg = mmc.ScatterConfig(pagesize=64)
    with pytest.raises(mmc.MallocMCError):
        mmc.compute_geometry(cfg, 64 * MiB)
    cfg = mmc.ScatterConfig(dataAlignment=24)
    with pytest.raises(mmc.MallocMCError):
        mmc.compute_geometry(cfg, 64 * MiB)


def test_cpp_programs_built():
    """The drop-in C++ headers compile into the example programs (built by build())."""
    import os
    d = os.path.join(REPO, "tests", "cpp", "_build")
    for prog in ("example_arrays", "example_slots", "verify_heap"):
        assert os.path.exists(os.path.join(d, prog)), prog
