"""XMallocSIMD on wave64: the coalesced chunk's layout against its closed form.

In the reference, collect() admits a lane when 0 < bytes < pagesize/32; the
leader allocates 16 + the sum of the group's bytes and writes the group's
thread count into the first word, and every lane's slice starts at chunk +
the offset an LDS atomic gave it (XMallocSIMD.hpp:118-176).  destroy() of a
slice decrements that header word, and only the slice that takes it from 1 to
0 releases the chunk (Scatter.hpp:924-940).  The wave64 form admits a lane
when 0 < bytes < pagesize/64, so that 16 + 64 slices fit a page, and assigns
offsets in lane order from a ballot scan.  So, for one full wave:
  * the group is the set C of lanes with 0 < pad16(bytes) < pagesize/64, the
    leader its lowest lane, the header word |C|;
  * lane l of C gets chunk + 16 + sum of pad16(bytes_j) over j in C, j < l;
  * the chunk's page is formatted for max(16 + sum, padded min chunk) bytes;
  * lanes outside C get allocations of their own (0 bytes: null);
  * freeing the slices one at a time, in any order, counts the header down
    and releases the chunk with the last one only.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MiB = 1024 ** 2


def pad16(b):
    return (b + 15) // 16 * 16


def _lane_sizes(P, seed):
    rng = np.random.default_rng(seed)
    lim = P // 64
    sizes = rng.integers(1, lim - 15, size=64)            # coalescible after padding
    sizes[rng.random(64) < 0.15] = 0                       # null lanes
    big = rng.random(64) < 0.15
    sizes[big] = rng.integers(lim, 4 * lim, size=int(big.sum()))  # too large to coalesce
    return sizes.astype(np.int32)


def _snapshot_words(snap_mod, a, byte_offset):
    s = snap_mod.take(a, payload=True)
    h = snap_mod.header(s)
    header_word = int(np.frombuffer(bytes(s), dtype=np.uint32, count=1, offset=h.off_payload + byte_offset)[0])
    return s, h, header_word


def _pte(s, h, page):
    cs, cnt, bits = np.frombuffer(bytes(s), dtype=np.uint32, count=3, offset=h.off_ptes + 12 * page).tolist()
    return cs, cnt, bits


@pytest.mark.parametrize("P,seed", [(4096, 1), (4096, 2), (16384, 3)])
def test_xmallocsimd_slice_layout_and_release(mmc, cuda, P, seed):
    from mallocmc_amd import snapshot as snap
    torch = cuda
    sizes = _lane_sizes(P, seed)
    padded = [pad16(int(b)) for b in sizes]
    C = [l for l in range(64) if 0 < padded[l] < P // 64]
    assert len(C) > 1
    total = sum(padded[l] for l in C)
    print(f"XMallocSIMD P={P} seed={seed}: {len(C)} coalesced lanes, {total} B", flush=True)

    cfg = mmc.ScatterConfig(pagesize=P, accessblocksize=1 * MiB, regionsize=16, resetfreedpages=True,
                            distribution=mmc.DIST_XMALLOCSIMD)
    a = mmc.ScatterAllocator(cfg, 2 * MiB)
    s0 = [a.getAvailableSlots(x) for x in (16, 256, P)]
    offs = a.malloc_batch(torch.from_numpy(sizes).cuda()).cpu().numpy()

    leader = C[0]
    base = int(offs[leader]) - 16
    prefix = 0
    for l in C:
        assert offs[l] == base + 16 + prefix, f"lane {l}"
        prefix += padded[l]
    for l in range(64):
        if padded[l] == 0:
            assert offs[l] == -1, f"lane {l}"
        elif l not in C:
            assert offs[l] >= 0 and (offs[l] + padded[l] <= base or offs[l] >= base + 16 + total), f"lane {l}"

    page = base // P
    s, h, hdr = _snapshot_words(snap, a, base)
    assert hdr == len(C)
    mc = pad16((P - 128) // 1024)
    cs, cnt, bits = _pte(s, h, page)
    assert cs == max(16 + total, mc) and (base - page * P) % cs == 0
    assert cnt == 1

    # slices freed one launch at a time, in a shuffled order
    order = np.random.default_rng(seed + 100).permutation(C)
    for k, l in enumerate(order):
        a.free_batch(torch.tensor([int(offs[l])], dtype=torch.int64, device="cuda"))
        s, h, hdr = _snapshot_words(snap, a, base)
        cs_k, cnt_k, _ = _pte(s, h, page)
        if k < len(C) - 1:
            assert hdr == len(C) - k - 1, f"after {k + 1} slice frees"
            assert (cs_k, cnt_k) == (cs, 1), f"chunk released early, after {k + 1} slice frees"
        else:
            assert hdr == 0
            assert (cs_k, cnt_k) == (0, 0), "the last slice free must release the chunk"
    others = [int(offs[l]) for l in range(64) if l not in C and offs[l] >= 0]
    a.free_batch(torch.tensor(others, dtype=torch.int64, device="cuda"))
    torch.cuda.synchronize()
    assert [a.getAvailableSlots(x) for x in (16, 256, P)] == s0
    a.close()
