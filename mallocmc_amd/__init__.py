"""MI355X-native ScatterAlloc device heap: Python host binding of libmallocmc_hip.so.

The host class mirrors ``mallocMC::Allocator`` (reference
``src/include/mallocMC/allocator.hpp:97-242``): construction reserves and
initialises the heap, ``getAvailableSlots`` counts free slots of a size,
``getHeapLocations`` returns the pool, ``destructiveResize`` rebuilds it and
``info`` prints the policy line-up.  Device-side ``malloc``/``free`` run in
batch launches (one lane per request) through the C ABI declared in
``include/mallocmc_hip.h``; C++/HIP callers include ``include/mallocMC`` and
call them from their own kernels.

PyTorch is only plumbing here (device buffers and the current stream).  There
is no CPU fallback: every method fails loudly when the HIP library is missing.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libmallocmc_hip.so")
NULL_OFFSET = -1  # MMC_NULL_OFFSET viewed as int64

DIST_NOOP = 0
DIST_XMALLOCSIMD = 1
FLAG_FIX_PAGESIZE_LEAK = 1

# Symbols the C ABI headers declare (include/mallocmc_hip.h, include/mallocmc_harness.h,
# include/mallocmc_batched.h).
EXPORTED_SYMBOLS = [
    "mmc_last_error", "mmc_version", "mmc_default_config", "mmc_compute_geometry", "mmc_create",
    "mmc_create_in_pool",
    "mmc_destroy", "mmc_destructive_resize", "mmc_heap_location", "mmc_get_config", "mmc_get_geometry",
    "mmc_info", "mmc_available_slots", "mmc_available_slots_async", "mmc_get_handle",
    "mmc_malloc_batch", "mmc_free_batch", "mmc_serial_stream",
    "mmc_write_pattern", "mmc_verify_pattern", "mmc_alloc_until_null", "mmc_bench_atomics",
    "mmc_bench_copy", "mmc_selftest_wave", "mmc_selftest_fastdiv", "mmc_selftest_pinned_geometry",
]
BATCHED_SYMBOLS = [
    "mmc_default_tuning", "mmc_wave_malloc_batch", "mmc_wave_free_batch", "mmc_multi_create",
    "mmc_multi_destroy", "mmc_multi_step", "mmc_multi_step_static", "mmc_wave_fill_until_null",
    # include/mallocmc_serve.h, bound in mallocmc_amd/serve.py
    "mmc_serve_create", "mmc_serve_destroy", "mmc_serve_load", "mmc_serve_step", "mmc_serve_get_stats",
    "mmc_serve_timestamps", "mmc_serve_stamp", "mmc_serve_stamp_count",
    # include/mallocmc_snapshot.h, bound in mallocmc_amd/snapshot.py
    "mmc_snapshot_bytes", "mmc_snapshot", "mmc_snapshot_verify",
]


class MmcTuning(ctypes.Structure):
    _fields_ = [
        ("aggregation", ctypes.c_uint32),
        ("xcd_slices", ctypes.c_uint32),
        ("persistent", ctypes.c_uint32),
        ("waves_per_cu", ctypes.c_uint32),
        ("class_split", ctypes.c_uint32),
        ("flat_single", ctypes.c_uint32),
        ("no_skip", ctypes.c_uint32),
        ("no_partition", ctypes.c_uint32),
        ("block_threads", ctypes.c_uint32),
        ("fresh_single", ctypes.c_uint32),
        ("lane_requests", ctypes.c_uint32),
    ]


@dataclasses.dataclass
class Tuning:
    """Wave-path tuning (mmc_tuning): aggregation width 1..64, XCD slices,
    persistent launch and its waves per CU, size classes per octave,
    whether pages are claimed one at a time (flat_single), whether
    the region scan walks every window (no_skip) and whether a block serves
    its requests in arrival order instead of partitioned by size
    (no_partition), the threads per block of the non-persistent wave
    malloc and free, of the non-persistent multi-heap step and of the fill
    (block_threads: 256, 512 or 1024; 0 = 1024), and whether unformatted
    pages of hierarchical chunk sizes are claimed one at a time
    (fresh_single), and the requests each lane of the partitioned wave
    malloc serves (lane_requests: 1, 2 or 4; 0 = 1; the free then serves
    two per lane)."""
    aggregation: int = 64
    xcd_slices: int = 0
    persistent: bool = False
    waves_per_cu: int = 8
    class_split: int = 0
    flat_single: bool = False
    no_skip: bool = False
    no_partition: bool = False
    block_threads: int = 0
    fresh_single: bool = False
    lane_requests: int = 0

    def to_c(self) -> MmcTuning:
        return MmcTuning(self.aggregation, self.xcd_slices, int(bool(self.persistent)), self.waves_per_cu,
                         self.class_split, int(bool(self.flat_single)), int(bool(self.no_skip)),
                         int(bool(self.no_partition)), self.block_threads, int(bool(self.fresh_single)),
                         self.lane_requests)


class MmcStepDesc(ctypes.Structure):
    _fields_ = [
        ("d_slot_heap", ctypes.c_void_p),
        ("d_tiles", ctypes.c_void_p),
        ("n_tiles", ctypes.c_uint32),
        ("n_slots", ctypes.c_uint32),
        ("d_size_table", ctypes.c_void_p),
        ("d_size_cdf", ctypes.c_void_p),
        ("n_sizes", ctypes.c_uint32),
        ("p_free_permille", ctypes.c_uint32),
        ("p_alloc_permille", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("d_offsets", ctypes.c_void_p),
        ("d_sizes", ctypes.c_void_p),
        ("d_step", ctypes.c_void_p),
        ("d_counters", ctypes.c_void_p),
        ("d_tile_cursor", ctypes.c_void_p),
    ]


class MmcConfig(ctypes.Structure):
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
        ("distribution", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 3),
    ]


class MmcGeometry(ctypes.Structure):
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
        ("sideMaskWords", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
    ]


@dataclasses.dataclass
class ScatterConfig:
    """ScatterConf::DefaultScatterConfig + DefaultScatterHashingParams
    (Scatter.hpp:56-85), Shrink's dataAlignment (Shrink.hpp:60-66; 1 = Noop
    alignment) and the distribution policy."""
    pagesize: int = 4096
    accessblocksize: int = 2 * 1024 ** 3
    regionsize: int = 16
    wastefactor: int = 2
    resetfreedpages: bool = False
    hashingK: int = 38183
    hashingDistMP: int = 17497
    hashingDistWP: int = 1
    hashingDistWPRel: int = 1
    dataAlignment: int = 16
    distribution: int = DIST_NOOP
    flags: int = 0

    def to_c(self) -> MmcConfig:
        c = MmcConfig()
        c.pagesize = self.pagesize
        c.regionsize = self.regionsize
        c.accessblocksize = self.accessblocksize
        c.wastefactor = self.wastefactor
        c.resetfreedpages = int(bool(self.resetfreedpages))
        c.hashingK = self.hashingK
        c.hashingDistMP = self.hashingDistMP
        c.hashingDistWP = self.hashingDistWP
        c.hashingDistWPRel = self.hashingDistWPRel
        c.dataAlignment = self.dataAlignment
        c.distribution = self.distribution
        c.flags = self.flags
        return c


class MallocMCError(RuntimeError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libmallocmc_hip.so (built in-tree by ``python -m mallocmc_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MallocMCError(f"{LIB_PATH} is missing: build it with `python -m mallocmc_amd.build`")
    # PyTorch-ROCm ships its own HIP runtime with the same soname
    # (libamdhip64.so.7).  Loading torch first makes the dynamic linker bind
    # this library to that one runtime instance; two HIP runtimes in one
    # process do not share the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "mmc_last_error": ([], ctypes.c_char_p),
        "mmc_version": ([], ctypes.c_char_p),
        "mmc_default_config": ([P(MmcConfig)], i32),
        "mmc_compute_geometry": ([P(MmcConfig), u64, P(MmcGeometry)], i32),
        "mmc_create": ([P(MmcConfig), u64, vp, P(vp)], i32),
        "mmc_create_in_pool": ([P(MmcConfig), vp, u64, vp, P(vp)], i32),
        "mmc_destroy": ([vp], i32),
        "mmc_destructive_resize": ([vp, u64, vp], i32),
        "mmc_heap_location": ([vp, P(vp), P(u64)], i32),
        "mmc_get_config": ([vp, P(MmcConfig)], i32),
        "mmc_get_geometry": ([vp, P(MmcGeometry)], i32),
        "mmc_info": ([vp, ctypes.c_char_p, u64], i32),
        "mmc_available_slots": ([vp, u64, vp, P(u32)], i32),
        "mmc_available_slots_async": ([vp, u64, vp, vp], i32),
        "mmc_get_handle": ([vp, vp, P(u64)], i32),
        "mmc_malloc_batch": ([vp, vp, vp, u64, vp], i32),
        "mmc_free_batch": ([vp, vp, u64, vp], i32),
        "mmc_serial_stream": ([vp, vp, vp, u64, vp], i32),
        "mmc_write_pattern": ([vp, vp, vp, u64, u64, vp], i32),
        "mmc_verify_pattern": ([vp, vp, vp, u64, u64, vp, vp], i32),
        "mmc_alloc_until_null": ([vp, u32, vp, u64, vp, u32, u32, vp], i32),
        "mmc_bench_atomics": ([vp, u64, u32, u32, u32, u32, vp, vp], i32),
        "mmc_bench_copy": ([vp, vp, u64, vp], i32),
        "mmc_selftest_wave": ([u32, u64, vp, vp], i32),
        "mmc_selftest_fastdiv": ([u64, u64, P(u64)], i32),
        "mmc_selftest_pinned_geometry": ([P(u32), P(u32)], i32),
        "mmc_default_tuning": ([P(MmcTuning)], i32),
        "mmc_wave_malloc_batch": ([vp, vp, vp, u64, P(MmcTuning), vp], i32),
        "mmc_wave_free_batch": ([vp, vp, u64, P(MmcTuning), vp], i32),
        "mmc_multi_create": ([P(vp), u32, P(vp)], i32),
        "mmc_multi_destroy": ([vp], i32),
        "mmc_multi_step": ([vp, P(MmcStepDesc), P(MmcTuning), vp], i32),
        "mmc_multi_step_static": ([P(vp), u32, P(MmcStepDesc), P(MmcTuning), vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().mmc_last_error()
        raise MallocMCError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _require_cuda(t, name):
    if not t.is_cuda:
        raise MallocMCError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise MallocMCError(f"{name} must be contiguous")


def compute_geometry(config: ScatterConfig, heap_bytes: int) -> MmcGeometry:
    g = MmcGeometry()
    c = config.to_c()
    _check(lib().mmc_compute_geometry(ctypes.byref(c), heap_bytes, ctypes.byref(g)), "mmc_compute_geometry")
    return g


class ScatterAllocator:
    """Host class of the heap (``mallocMC::Allocator``, allocator.hpp:97-242)."""

    def __init__(self, config: ScatterConfig | None = None, heap_bytes: int = 8 * 1024 * 1024, stream=None):
        self.config = config or ScatterConfig()
        self._h = ctypes.c_void_p()
        c = self.config.to_c()
        _check(lib().mmc_create(ctypes.byref(c), heap_bytes, _stream_ptr(stream), ctypes.byref(self._h)),
               "mmc_create")

    # -------------------------------------------------------------- lifecycle
    def close(self):
        if self._h:
            lib().mmc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._h:
            raise MallocMCError("allocator is closed")
        return self._h

    def destructiveResize(self, heap_bytes: int, stream=None):
        _check(lib().mmc_destructive_resize(self.handle, heap_bytes, _stream_ptr(stream)), "destructiveResize")

    def getHeapLocations(self):
        p, n = ctypes.c_void_p(), ctypes.c_uint64()
        _check(lib().mmc_heap_location(self.handle, ctypes.byref(p), ctypes.byref(n)), "getHeapLocations")
        return [(p.value or 0, n.value)]

    def geometry(self) -> MmcGeometry:
        g = MmcGeometry()
        _check(lib().mmc_get_geometry(self.handle, ctypes.byref(g)), "mmc_get_geometry")
        return g

    def info(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        _check(lib().mmc_info(self.handle, buf, len(buf)), "info")
        return buf.value.decode()

    def getAvailableSlots(self, slot_size: int, stream=None) -> int:
        out = ctypes.c_uint32()
        _check(lib().mmc_available_slots(self.handle, slot_size, _stream_ptr(stream), ctypes.byref(out)),
               "getAvailableSlots")
        return out.value

    def available_slots_async(self, slot_size: int, out_tensor, stream=None):
        _require_cuda(out_tensor, "out_tensor")
        _check(lib().mmc_available_slots_async(self.handle, slot_size, _ptr(out_tensor), _stream_ptr(stream)),
               "available_slots_async")

    # --------------------------------------------------------- device batches
    def malloc_batch(self, sizes, offsets=None, stream=None):
        """Lane i mallocs sizes[i] (int32 device tensor, read as uint32);
        returns int64 pool offsets, -1 for null."""
        import torch
        _require_cuda(sizes, "sizes")
        if offsets is None:
            offsets = torch.empty(sizes.numel(), dtype=torch.int64, device=sizes.device)
        _require_cuda(offsets, "offsets")
        _check(lib().mmc_malloc_batch(self.handle, _ptr(sizes), _ptr(offsets), sizes.numel(), _stream_ptr(stream)),
               "malloc_batch")
        return offsets

    def free_batch(self, offsets, stream=None):
        _require_cuda(offsets, "offsets")
        _check(lib().mmc_free_batch(self.handle, _ptr(offsets), offsets.numel(), _stream_ptr(stream)), "free_batch")

    def wave_malloc_batch(self, sizes, offsets=None, tuning: Tuning | None = None, stream=None):
        """Wave-cooperative malloc of sizes[i] (same contract as malloc_batch)."""
        import torch
        _require_cuda(sizes, "sizes")
        if offsets is None:
            offsets = torch.empty(sizes.numel(), dtype=torch.int64, device=sizes.device)
        t = (tuning or Tuning()).to_c()
        _check(lib().mmc_wave_malloc_batch(self.handle, _ptr(sizes), _ptr(offsets), sizes.numel(), ctypes.byref(t),
                                           _stream_ptr(stream)), "wave_malloc_batch")
        return offsets

    def wave_free_batch(self, offsets, tuning: Tuning | None = None, stream=None):
        _require_cuda(offsets, "offsets")
        t = (tuning or Tuning()).to_c()
        _check(lib().mmc_wave_free_batch(self.handle, _ptr(offsets), offsets.numel(), ctypes.byref(t),
                                         _stream_ptr(stream)), "wave_free_batch")

    def serial_stream(self, ops, stream=None):
        """Serialized single-lane stream with the CPU backend's ids (see C ABI)."""
        import torch
        _require_cuda(ops, "ops")
        out = torch.empty(ops.numel(), dtype=torch.int64, device=ops.device)
        _check(lib().mmc_serial_stream(self.handle, _ptr(ops), _ptr(out), ops.numel(), _stream_ptr(stream)),
               "serial_stream")
        return out

    # ------------------------------------------------------ verify_heap harness
    def write_pattern(self, offsets, sizes, salt: int = 0, stream=None):
        _check(lib().mmc_write_pattern(self.handle, _ptr(offsets), _ptr(sizes), offsets.numel(), salt,
                                       _stream_ptr(stream)), "write_pattern")

    def verify_pattern(self, offsets, sizes, salt: int = 0, stream=None):
        """Returns dict(mismatches, misaligned, outside, nulls, checked)."""
        import torch
        res = torch.zeros(5, dtype=torch.int64, device=offsets.device)
        _check(lib().mmc_verify_pattern(self.handle, _ptr(offsets), _ptr(sizes), offsets.numel(), salt, _ptr(res),
                                        _stream_ptr(stream)), "verify_pattern")
        v = res.cpu().tolist()
        return dict(mismatches=v[0], misaligned=v[1], outside=v[2], nulls=v[3], checked=v[4])

    def alloc_until_null(self, size: int, capacity: int, blocks: int = 64, threads: int = 128, stream=None):
        import torch
        offs = torch.full((capacity,), -1, dtype=torch.int64, device="cuda")
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        _check(lib().mmc_alloc_until_null(self.handle, size, _ptr(offs), capacity, _ptr(cnt), blocks, threads,
                                          _stream_ptr(stream)), "alloc_until_null")
        n = int(cnt.item())
        return offs[: min(n, capacity)], n


class MultiHeap:
    """Many independent heaps served by one launch per step (mmc_multi_*):
    persistent waves pulling 64-slot tiles (Tuning(persistent=True), the
    default here) or one workgroup per block_threads/64 tiles
    (persistent=False).  The
    heaps stay owned by their ScatterAllocator objects."""

    def __init__(self, allocators):
        self.allocators = list(allocators)
        arr = (ctypes.c_void_p * len(self.allocators))(*[a.handle.value for a in self.allocators])
        self._m = ctypes.c_void_p()
        _check(lib().mmc_multi_create(arr, len(self.allocators), ctypes.byref(self._m)), "mmc_multi_create")

    def close(self):
        if self._m:
            lib().mmc_multi_destroy(self._m)
            self._m = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, desc: MmcStepDesc, tuning: Tuning | None = None, stream=None):
        t = (tuning or Tuning(persistent=True)).to_c()
        _check(lib().mmc_multi_step(self._m, ctypes.byref(desc), ctypes.byref(t), _stream_ptr(stream)),
               "mmc_multi_step")


STATIC_MAX_HEAPS = 256  # MMC_STATIC_MAX_HEAPS


def multi_step_static(allocators, desc: MmcStepDesc, tuning: Tuning | None = None, stream=None):
    """One multi-heap step over `allocators` that allocates no device memory
    (mmc_multi_step_static): the heap table and tile cursor live in the
    library's fixed per-device array (at most STATIC_MAX_HEAPS heaps), so
    desc.d_tile_cursor is not used.  Calls that overlap on one device race."""
    heaps = (ctypes.c_void_p * len(allocators))(*[a.handle.value for a in allocators])
    t = (tuning or Tuning(persistent=True)).to_c()
    _check(lib().mmc_multi_step_static(heaps, len(allocators), ctypes.byref(desc), ctypes.byref(t),
                                       _stream_ptr(stream)), "mmc_multi_step_static")


def default_config() -> ScatterConfig:
    return ScatterConfig()
