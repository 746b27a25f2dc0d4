"""ctypes binding of the CPU oracle (scatter_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker / reported baseline.
The oracle restates mallocMC's Scatter policy on the reference's CPU backend
(``src/include/mallocMC/creationPolicies/Scatter.hpp``); see scatter_oracle.h
for the execution model and the documented divergences D1-D3.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libscatter_oracle.so")

ORC_NULL = (1 << 64) - 1
ATOMIC_LOCKFREE = 0
ATOMIC_MUTEX16 = 1


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("pagesize", ctypes.c_uint32),
        ("regionsize", ctypes.c_uint32),
        ("accessblocksize", ctypes.c_uint64),
        ("wastefactor", ctypes.c_uint32),
        ("resetfreedpages", ctypes.c_uint32),
        ("hashingK", ctypes.c_uint32),
        ("hashingDistMP", ctypes.c_uint32),
        ("hashingDistWP", ctypes.c_uint32),
        ("hashingDistWPRel", ctypes.c_uint32),
        ("dataAlignment", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class OrcGeometry(ctypes.Structure):
    _fields_ = [
        ("numpages", ctypes.c_uint32),
        ("numregions", ctypes.c_uint32),
        ("accessblocks", ctypes.c_uint32),
        ("pagesperblock", ctypes.c_uint32),
        ("hierarchyThreshold", ctypes.c_uint32),
        ("minChunkSize", ctypes.c_uint32),
        ("maxOnPageMasks", ctypes.c_uint32),
        ("clampedBlocks", ctypes.c_uint32),
        ("ptesOffset", ctypes.c_uint64),
        ("regionsOffset", ctypes.c_uint64),
        ("usedBytes", ctypes.c_uint64),
    ]


class OrcBenchResult(ctypes.Structure):
    _fields_ = [
        ("t_malloc", ctypes.c_double),
        ("t_write_verify", ctypes.c_double),
        ("t_free", ctypes.c_double),
        ("nulls", ctypes.c_uint64),
        ("mismatches", ctypes.c_uint64),
        ("threads_used", ctypes.c_int),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_libs = {}


def _load(path: str):
    L = _libs.get(path)
    if L is not None:
        return L
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    L.orc_default_config.argtypes = [P(OrcConfig)]
    L.orc_create.argtypes = [P(OrcConfig), ctypes.c_uint64, P(ctypes.c_char_p)]
    L.orc_create.restype = ctypes.c_void_p
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_get_geometry.argtypes = [ctypes.c_void_p, P(OrcGeometry)]
    L.orc_compute_geometry.argtypes = [P(OrcConfig), ctypes.c_uint64, P(OrcGeometry)]
    L.orc_compute_geometry.restype = ctypes.c_int
    L.orc_set_atomic_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_malloc.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.orc_malloc.restype = ctypes.c_uint64
    L.orc_free.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.orc_available_slots.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.orc_available_slots.restype = ctypes.c_uint32
    L.orc_run_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    L.orc_read_pte.argtypes = [ctypes.c_void_p, ctypes.c_uint32] + [P(ctypes.c_uint32)] * 3
    L.orc_read_region.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.orc_read_region.restype = ctypes.c_uint32
    L.orc_read_scalars.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32), P(ctypes.c_uint32)]
    L.orc_pool.argtypes = [ctypes.c_void_p]
    L.orc_pool.restype = ctypes.c_void_p
    L.orc_bench_concurrent.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, P(OrcBenchResult)]
    L.orc_bench_concurrent.restype = ctypes.c_int
    _libs[path] = L
    return L


def lib():
    """The faithful oracle build."""
    return _load(_LIB_PATH)


def mutant_lib(n: int):
    """Test-only mutant build n (scatter_oracle.c, ORC_MUTANT): one stateful
    heuristic broken, for showing that the known-answer tests catch it."""
    return _load(os.path.join(_HERE, "_build", f"libscatter_oracle_mut{n}.so"))


def make_config(pagesize=4096, regionsize=16, accessblocksize=2 * 1024 ** 3, wastefactor=2,
                resetfreedpages=False, hashingK=38183, hashingDistMP=17497, hashingDistWP=1,
                hashingDistWPRel=1, dataAlignment=16) -> OrcConfig:
    return OrcConfig(pagesize, regionsize, accessblocksize, wastefactor, int(bool(resetfreedpages)),
                     hashingK, hashingDistMP, hashingDistWP, hashingDistWPRel, dataAlignment, 0)


def geometry(cfg: OrcConfig, memsize: int) -> OrcGeometry:
    g = OrcGeometry()
    rc = lib().orc_compute_geometry(ctypes.byref(cfg), memsize, ctypes.byref(g))
    if rc != 0:
        raise ValueError("invalid configuration")
    return g


class OracleHeap:
    """One CPU heap; methods mirror DeviceAllocator::malloc/free and
    Allocator::getAvailableSlots on the reference's CPU backend."""

    def __init__(self, cfg: OrcConfig, memsize: int, library=None):
        err = ctypes.c_char_p()
        self._cfg = cfg
        self._L = library if library is not None else lib()
        self._h = self._L.orc_create(ctypes.byref(cfg), memsize, ctypes.byref(err))
        if not self._h:
            raise ValueError((err.value or b"orc_create failed").decode())
        self.memsize = memsize

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def geometry(self) -> OrcGeometry:
        g = OrcGeometry()
        self._L.orc_get_geometry(self._h, ctypes.byref(g))
        return g

    def set_atomic_mode(self, mode: int):
        self._L.orc_set_atomic_mode(self._h, mode)

    def malloc(self, nbytes: int):
        off = self._L.orc_malloc(self._h, nbytes)
        return None if off == ORC_NULL else off

    def free(self, offset):
        self._L.orc_free(self._h, ORC_NULL if offset is None else offset)

    def available_slots(self, size: int) -> int:
        return self._L.orc_available_slots(self._h, size)

    def run_stream(self, ops) -> np.ndarray:
        ops = np.ascontiguousarray(ops, dtype=np.int64)
        out = np.zeros(len(ops), dtype=np.uint64)
        self._L.orc_run_stream(self._h, ops.ctypes.data, len(ops), out.ctypes.data)
        return out

    def pte(self, page: int):
        a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        self._L.orc_read_pte(self._h, page, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def region(self, r: int) -> int:
        return self._L.orc_read_region(self._h, r)

    def scalars(self):
        a, b = ctypes.c_uint32(), ctypes.c_uint32()
        self._L.orc_read_scalars(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def pool_bytes(self, offset: int, n: int) -> bytes:
        base = self._L.orc_pool(self._h)
        return ctypes.string_at(base + offset, n)

    def bench_concurrent(self, sizes, nthreads: int = 0):
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        offs = np.zeros(len(sizes), dtype=np.uint64)
        res = OrcBenchResult()
        self._L.orc_bench_concurrent(self._h, sizes.ctypes.data, len(sizes), nthreads, offs.ctypes.data,
                                   ctypes.byref(res))
        return res, offs
