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
        if f in ("sideMaskWords", "reserved0"):  # side mask array: no reference counterpart
            continue
        assert getattr(g, f) == getattr(og, f), f


@pytest.mark.parametrize("pagesize", [128, 256, 1000, 1024, 4096, 8192, 16384, 65536, 1 << 20])
@pytest.mark.parametrize("align", [1, 4, 16, 64])
def test_side_mask_words_cover_every_hierarchical_chunk(mmc, oracle_mod, pagesize, align):
    """Brute force over every chunk size a page can take: the segments it
    needs (full segments plus a partial one, Scatter.hpp:365-383, 494-510)
    never exceed the side array's words per page, and the smallest chunk
    needs all of them or one fewer."""
    cfg = mmc.ScatterConfig(pagesize=pagesize, accessblocksize=64 * MiB, regionsize=16, dataAlignment=align)
    g = mmc.compute_geometry(cfg, 64 * MiB)
    og = oracle_mod.geometry(oracle_mod.make_config(pagesize=pagesize, accessblocksize=64 * MiB, regionsize=16), 64 * MiB)
    ht, mc = og.hierarchyThreshold, og.minChunkSize
    assert 1 <= g.sideMaskWords <= 32
    cmin = max(-(-mc // align) * align, align)
    most = 0
    for c in range(cmin, ht + 1, align):
        seg = 32 * c + 4
        full = min(32, pagesize // seg)
        add = 0 if full == 32 else max(0, pagesize - full * seg - 4) // c
        most = max(most, full + (1 if add > 0 else 0))
    assert most <= g.sideMaskWords
    if most:
        assert g.sideMaskWords - most <= 1


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
    for prog in ("example_arrays", "example_slots", "verify_heap", "thread_malloc", "custom_policies",
                 "udiv24_exhaustive"):
        assert os.path.exists(os.path.join(d, prog)), prog


def test_serve_payload_word_matches_header(tmp_path):
    """The host closed form of the serving payload (mallocmc_amd/serve.py)
    equals mmc_serve_word of include/mallocmc_serve.h compiled with gcc."""
    import subprocess
    import numpy as np
    from mallocmc_amd import serve
    src = tmp_path / "w.c"
    src.write_text('#include "mallocmc_serve.h"\nunsigned w(unsigned r, unsigned t, unsigned x)'
                   '{ return mmc_serve_word(r, t, x); }\n')
    so = tmp_path / "w.so"
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-I", os.path.join(REPO, "include"), str(src), "-o",
                    str(so)], check=True)
    L = ctypes.CDLL(str(so))
    L.w.restype = ctypes.c_uint32
    L.w.argtypes = [ctypes.c_uint32] * 3
    rng = np.random.default_rng(0)
    for _ in range(200):
        r, t, x = (int(v) for v in rng.integers(0, 2 ** 32, size=3, dtype=np.uint64))
        got = int(serve.payload_words(r, np.array([t], np.uint64), np.array([x], np.uint64))[0])
        assert got == L.w(r, t, x)


def test_serve_checksum_closed_form_small():
    """expected_checksum against a direct loop for tiny requests."""
    import numpy as np
    from mallocmc_amd import serve
    for rid, length in ((0, 1), (7, 3), (123456, 5)):
        acc = 0
        for k in range(length * 64):
            x = int(serve.payload_words(rid, np.array([k // 64], np.uint64), np.array([k % 64], np.uint64))[0])
            acc = (acc + x * (2 * k + 1)) % (1 << 64)
        assert serve.expected_checksum(rid, length) == acc


def test_serve_trace_shape():
    import numpy as np
    from mallocmc_amd import serve
    t = serve.make_trace(seed=5)
    assert (np.diff(t.arrival.astype(np.int64)) >= 0).all()
    assert t.length.min() >= 1 and t.length.max() <= 8192
    assert abs(len(t.arrival) / 30.0 - 2000) < 100  # Poisson 2000/s over 30 s


def test_pinned_geometries_accept_their_heaps_only(mmc):
    """The batch kernels built for fixed configurations (and for the default
    tuning) run only on heaps whose view holds exactly their compile-time
    constants, and only with the default tuning."""
    pinned, wrong = ctypes.c_uint32(0), ctypes.c_uint32(0)
    assert mmc.lib().mmc_selftest_pinned_geometry(ctypes.byref(pinned), ctypes.byref(wrong)) == 0
    assert pinned.value == 5
    assert wrong.value == 0


def test_fastdiv_matches_division(mmc):
    """The heap view's division by invariant divisors (multiply-high and
    shifts) equals `/` and `%` on edge cases and 2,000,000 seeded pairs
    (host run of the same header code the kernels use)."""
    import ctypes
    bad = ctypes.c_uint64(123)
    assert mmc.lib().mmc_selftest_fastdiv(7, 1_000_000, ctypes.byref(bad)) == 0
    assert bad.value == 0


def test_serve_checksums_vectorized_equal_closed_form():
    """The chunked torch form used to check every served request equals the
    per-request closed form, across chunk boundaries and at the length cap."""
    import numpy as np
    from mallocmc_amd import serve
    rng = np.random.default_rng(2)
    ids = rng.integers(0, 2 ** 20, 12)
    lens = rng.integers(1, 200, 12)
    lens[0], lens[1] = 8192, 1
    got = serve.expected_checksums(ids, lens, device="cpu", chunk_terms=1 << 14)
    want = np.array([serve.expected_checksum(int(i), int(n)) for i, n in zip(ids, lens)], dtype=np.uint64)
    np.testing.assert_array_equal(got, want)


def test_serve_trace_validation():
    """Zero or over-long lengths and decreasing arrivals are rejected before a
    trace reaches the device (mmc_serve_load checks them again there)."""
    import numpy as np
    from mallocmc_amd import serve
    ok = serve.Trace(np.array([0, 0, 1], np.uint32), np.array([1, 8192, 5], np.uint32), np.arange(3, dtype=np.uint32))
    serve.validate_trace(ok, 8192)
    for arrival, length, msg in (([0, 0, 1], [1, 0, 5], "request 1: length 0"),
                                 ([0, 0, 1], [1, 8193, 5], "request 1: length 8193"),
                                 ([0, 2, 1], [1, 1, 1], "request 2: arrival")):
        t = serve.Trace(np.array(arrival, np.uint32), np.array(length, np.uint32), np.arange(3, dtype=np.uint32))
        with pytest.raises(ValueError, match=msg):
            serve.validate_trace(t, 8192)
