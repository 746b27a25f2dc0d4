"""Builds a heap snapshot (include/mallocmc_snapshot.h format) from the
oracle's state, so that the host checker can be exercised on states the
reference protocol reaches (tests only).  The oracle keeps the reference's
pool layout with on-page masks; they are copied into the side-mask section."""
import ctypes

import numpy as np


def _align8(x):
    return (x + 7) & ~7


def snapshot_from_oracle(mmc, snapmod, h, scfg, memsize):
    g = mmc.compute_geometry(scfg, memsize)
    P, n = scfg.pagesize, g.numpages
    hdr = snapmod.MmcSnapshotHeader()
    hdr.magic = b"MMCSNAP1"
    hdr.version = 1
    hdr.flags = 0
    hdr.cfg = scfg.to_c()
    hdr.geom = g
    hdr.memsize = memsize
    ffpb, ffb = h.scalars()
    hdr.firstFreePageBased, hdr.firstfreeblock, hdr.pagebasedMutex = ffpb, ffb, 0
    off_ptes = _align8(ctypes.sizeof(hdr))
    off_regions = _align8(off_ptes + 12 * n)
    off_masks = _align8(off_regions + 4 * g.numregions)
    total = _align8(off_masks + 4 * g.sideMaskWords * n)
    hdr.off_ptes, hdr.off_regions, hdr.off_masks, hdr.off_payload, hdr.total_bytes = (
        off_ptes, off_regions, off_masks, 0, total)
    buf = bytearray(total)
    buf[: ctypes.sizeof(hdr)] = bytes(hdr)
    ptes = h.pool_bytes(g.ptesOffset, 12 * n)
    buf[off_ptes: off_ptes + 12 * n] = ptes
    buf[off_regions: off_regions + 4 * g.numregions] = h.pool_bytes(g.regionsOffset, 4 * g.numregions)
    pte = np.frombuffer(ptes, dtype=np.uint32).reshape(n, 3)
    masks = np.zeros((n, g.sideMaskWords), dtype=np.uint32)
    ht = g.hierarchyThreshold
    for p in np.nonzero((pte[:, 0] > 0) & (pte[:, 0] <= ht))[0]:
        cs = int(pte[p, 0])
        seg = 32 * cs + 4
        full = min(32, P // seg)
        add = 0 if full == 32 else min(31, max(0, P - full * seg - 4) // cs)
        nm = full + (1 if add > 0 else 0)
        words = np.frombuffer(h.pool_bytes(int(p) * P + P - 4 * nm, 4 * nm), dtype=np.uint32)
        masks[p, :nm] = words
    buf[off_masks: off_masks + masks.nbytes] = masks.tobytes()
    return buf
