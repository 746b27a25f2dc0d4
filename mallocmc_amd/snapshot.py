"""Heap snapshots and offline verification (include/mallocmc_snapshot.h).

A snapshot is a relocatable host copy of a quiescent heap's metadata (PTEs,
region fill levels, segment masks, heap scalars; optionally the payload).
`verify` checks it on the host: per-page invariants, multi-page runs, and,
given the caller's live allocations, that each is marked allocated, fits its
chunk and is not shared.
"""
from __future__ import annotations

import ctypes

import numpy as np

import mallocmc_amd as mmc

SNAPSHOT_PAYLOAD = 1
VERIFY_LIVE_COMPLETE = 1
VERIFY_LIVE_INTERIOR = 2
SNAPSHOT_SYMBOLS = ["mmc_snapshot_bytes", "mmc_snapshot", "mmc_snapshot_verify"]


class MmcSnapshotHeader(ctypes.Structure):
    _fields_ = [
        ("magic", ctypes.c_char * 8),
        ("version", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("cfg", mmc.MmcConfig),
        ("geom", mmc.MmcGeometry),
        ("memsize", ctypes.c_uint64),
        ("firstFreePageBased", ctypes.c_uint32),
        ("firstfreeblock", ctypes.c_uint32),
        ("pagebasedMutex", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("off_ptes", ctypes.c_uint64),
        ("off_regions", ctypes.c_uint64),
        ("off_masks", ctypes.c_uint64),
        ("off_payload", ctypes.c_uint64),
        ("total_bytes", ctypes.c_uint64),
    ]


class MmcVerifyReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "pages", "free_pages", "chunked_pages", "run_pages", "runs", "used_chunks", "empty_formatted",
        "stale_hints", "live_checked", "errors", "first_error_page")] + [
        ("first_error_code", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("message", ctypes.c_char * 160)]


def _bind():
    L = mmc.lib()
    if not getattr(L, "_snap_bound", False):
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.mmc_snapshot_bytes.argtypes = [vp, u32, ctypes.POINTER(u64)]
        L.mmc_snapshot.argtypes = [vp, vp, u64, u32, vp]
        L.mmc_snapshot_verify.argtypes = [vp, u64, vp, vp, u64, u32, ctypes.POINTER(MmcVerifyReport)]
        for n in SNAPSHOT_SYMBOLS:
            getattr(L, n).restype = ctypes.c_int
        L._snap_bound = True
    return L


def take(alloc: mmc.ScatterAllocator, payload: bool = False, stream=None) -> bytearray:
    """Copies the heap's state to a host buffer after the work queued on
    `stream`; the heap must be quiescent."""
    L = _bind()
    flags = SNAPSHOT_PAYLOAD if payload else 0
    n = ctypes.c_uint64()
    mmc._check(L.mmc_snapshot_bytes(alloc.handle, flags, ctypes.byref(n)), "mmc_snapshot_bytes")
    buf = bytearray(n.value)
    cbuf = (ctypes.c_char * len(buf)).from_buffer(buf)
    mmc._check(L.mmc_snapshot(alloc.handle, ctypes.addressof(cbuf), len(buf), flags, mmc._stream_ptr(stream)),
               "mmc_snapshot")
    return buf


def header(snap) -> MmcSnapshotHeader:
    return MmcSnapshotHeader.from_buffer_copy(bytes(snap[: ctypes.sizeof(MmcSnapshotHeader)]))


def save(path: str, snap) -> None:
    with open(path, "wb") as f:
        f.write(bytes(snap))


def load(path: str) -> bytearray:
    with open(path, "rb") as f:
        return bytearray(f.read())


def _host_array(x, dtype):
    if x is None:
        return None
    if hasattr(x, "detach"):  # a torch tensor, possibly on the GPU
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x).astype(dtype, copy=False))


def verify(snap, offsets=None, sizes=None, complete: bool = False, interior: bool = False) -> dict:
    """Checks a snapshot; offsets/sizes are the live allocations (pool
    offsets, bytes; entries < 0 are skipped)."""
    L = _bind()
    flags = (VERIFY_LIVE_COMPLETE if complete else 0) | (VERIFY_LIVE_INTERIOR if interior else 0)
    offs = _host_array(offsets, np.int64)
    szs = _host_array(sizes, np.uint32)
    if offs is not None:
        keep = offs >= 0
        offs = np.ascontiguousarray(offs[keep].astype(np.uint64))
        szs = np.ascontiguousarray(szs[keep]) if szs is not None else None
    buf = bytes(snap)
    rep = MmcVerifyReport()
    mmc._check(L.mmc_snapshot_verify(
        ctypes.c_char_p(buf), len(buf),
        offs.ctypes.data if offs is not None and len(offs) else None,
        szs.ctypes.data if szs is not None and len(offs) else None,
        len(offs) if offs is not None else 0, flags, ctypes.byref(rep)), "mmc_snapshot_verify")
    out = {n: getattr(rep, n) for n, _ in MmcVerifyReport._fields_ if n not in ("reserved", "message")}
    out["message"] = rep.message.decode(errors="replace")
    out["ok"] = rep.errors == 0
    return out
