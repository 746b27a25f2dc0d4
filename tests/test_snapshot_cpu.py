"""Host checker of heap snapshots (include/mallocmc_snapshot.h) on states the
reference protocol reaches: the oracle runs seeded request streams, its pool
is turned into a snapshot, and the checker must accept it with the complete
live list -- and reject each planted corruption with the right code."""
import ctypes

import numpy as np
import pytest

from snapshot_oracle import snapshot_from_oracle

MiB = 1024 ** 2


@pytest.fixture(scope="module")
def snapmod(mmc):
    from mallocmc_amd import snapshot
    return snapshot


def _run(oracle_mod, P, reset, seed, n_ops=3000, big=True, heap=8 * MiB):
    cfg = oracle_mod.make_config(pagesize=P, regionsize=16, accessblocksize=2 * MiB, resetfreedpages=int(reset))
    h = oracle_mod.OracleHeap(cfg, heap)
    rng = np.random.default_rng(seed)
    live = {}
    for _ in range(n_ops):
        if live and rng.random() < 0.45:
            off = list(live)[rng.integers(len(live))]
            h.free(off)
            del live[off]
        else:
            r = rng.random()
            if big and r < 0.03:
                b = int(rng.integers(P + 1, 3 * P))          # multi-page run
            elif r < 0.5:
                b = int(rng.integers(1, 200))                 # hierarchical chunks
            else:
                b = int(rng.integers(200, P - 16))            # flat chunks (never exactly P)
            off = h.malloc(b)
            if off is not None:
                live[off] = b
    return h, live


def _scfg(mmc, P, reset):
    return mmc.ScatterConfig(pagesize=P, regionsize=16, accessblocksize=2 * MiB, resetfreedpages=reset)


@pytest.mark.parametrize("P,reset,seed", [(4096, True, 1), (4096, False, 2), (16384, True, 3), (1024, True, 4)])
def test_reference_states_verify(mmc, oracle_mod, snapmod, P, reset, seed):
    h, live = _run(oracle_mod, P, reset, seed)
    snap = snapshot_from_oracle(mmc, snapmod, h, _scfg(mmc, P, reset), 8 * MiB)
    offs = np.array(list(live.keys()), dtype=np.int64)
    sizes = np.array([(b + 15) // 16 * 16 for b in live.values()], dtype=np.uint32)
    r = snapmod.verify(snap, offs, sizes, complete=True)
    assert r["ok"], r
    assert r["live_checked"] == len(live)
    assert r["chunked_pages"] > 0 and r["used_chunks"] > 0
    if reset:
        assert r["empty_formatted"] == 0, r  # serialized: the last free resets its page
    h.close()


def _corrupt(snap, off, value):
    b = bytearray(snap)
    b[off: off + 4] = np.uint32(value).tobytes()
    return b


def test_planted_corruptions_are_found(mmc, oracle_mod, snapmod):
    P, reset = 4096, True
    h, live = _run(oracle_mod, P, reset, 7)
    snap = snapshot_from_oracle(mmc, snapmod, h, _scfg(mmc, P, reset), 8 * MiB)
    hdr = snapmod.header(snap)
    g = hdr.geom
    offs = np.array(list(live.keys()), dtype=np.int64)
    sizes = np.array([(b + 15) // 16 * 16 for b in live.values()], dtype=np.uint32)
    assert snapmod.verify(snap, offs, sizes, complete=True)["ok"]
    pte = np.frombuffer(bytes(snap[hdr.off_ptes: hdr.off_ptes + 12 * g.numpages]), dtype=np.uint32).reshape(-1, 3)
    hier = np.nonzero((pte[:, 0] > 0) & (pte[:, 0] <= g.hierarchyThreshold) & (pte[:, 1] > 0))[0]
    flat = np.nonzero((pte[:, 0] > g.hierarchyThreshold) & (pte[:, 0] <= P) & (pte[:, 1] > 0))[0]
    free = np.nonzero(pte[:, 0] == 0)[0]
    runs = np.nonzero(pte[:, 0] > P)[0]
    assert len(hier) and len(flat) and len(free) and len(runs)

    def code(s, o=offs, z=sizes, complete=True):
        r = snapmod.verify(s, o, z, complete=complete)
        assert not r["ok"]
        return r["first_error_code"]

    p = int(free[0])  # a free page with a count
    assert code(_corrupt(snap, hdr.off_ptes + 12 * p + 4, 3)) == 1
    p = int(flat[0])  # a flat page's count off by one
    assert code(_corrupt(snap, hdr.off_ptes + 12 * p + 4, int(pte[p, 1]) + 1)) == 3
    # a flat page's bit beyond its capacity
    cap = min(P // int(pte[p, 0]), 32)
    if cap < 32:
        assert code(_corrupt(snap, hdr.off_ptes + 12 * p + 8, int(pte[p, 2]) | (1 << 31))) == 2
    p = int(hier[0])  # a hierarchical page's first mask word cleared: count no longer matches
    moff = hdr.off_masks + 4 * g.sideMaskWords * p
    assert code(_corrupt(snap, moff, 0)) in (3, 7)
    p = int(runs[0])  # a run's second page with another chunk size
    k = -(-int(pte[p, 0]) // P)
    if k > 1:
        assert code(_corrupt(snap, hdr.off_ptes + 12 * (p + 1), int(pte[p, 0]) + 16)) == 4
    # a live allocation listed twice
    assert code(snap, np.concatenate([offs, offs[:1]]), np.concatenate([sizes, sizes[:1]])) == 8
    # a live allocation larger than its chunk
    z = sizes.copy()
    i = int(np.nonzero(offs // P == int(flat[0]))[0][0])
    z[i] = int(pte[int(flat[0]), 0]) + 16
    assert code(snap, offs, z) == 7
    # an allocation missing from a complete list
    assert code(snap, offs[1:], sizes[1:]) == 9
    # truncated or foreign bytes
    assert code(snap[:100]) == 10
    bad = bytearray(snap)
    bad[0:8] = b"NOTASNAP"
    assert code(bad) == 10
    h.close()


def test_pagesize_chunk_leak_is_reported(mmc, oracle_mod, snapmod):
    """The reference releases a chunk of exactly pagesize bytes through the
    page-based path (Scatter.hpp:906-908 vs 942-945): its page keeps a count
    with no live chunk, which the checker reports (code 3 or 9)."""
    P = 4096
    cfg = oracle_mod.make_config(pagesize=P, regionsize=16, accessblocksize=2 * MiB, resetfreedpages=1)
    h = oracle_mod.OracleHeap(cfg, 8 * MiB)
    off = h.malloc(P)
    h.free(off)
    snap = snapshot_from_oracle(mmc, snapmod, h, _scfg(mmc, P, True), 8 * MiB)
    r = snapmod.verify(snap, np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.uint32), complete=True)
    assert not r["ok"] and r["first_error_code"] in (1, 3, 9), r
    h.close()
