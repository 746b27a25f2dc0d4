"""Serving loop driver: Poisson arrivals, growing paged buffers,
admission by free heap bytes, p50/p99 step latency.

The loop itself runs on the device (include/mallocmc_serve.h); this module
builds the seeded request trace, captures one step in a hipGraph, replays it
for every simulated step with HIP events around each replay, and checks
request checksums against their closed form.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

import mallocmc_amd as mmc

GiB = 1024 ** 3
MiB = 1024 ** 2
M32 = np.uint64(0xFFFFFFFF)


class MmcServeConfig(ctypes.Structure):
    _fields_ = [
        ("max_slots", ctypes.c_uint32),
        ("block_bytes", ctypes.c_uint32),
        ("tokens_per_block", ctypes.c_uint32),
        ("token_bytes", ctypes.c_uint32),
        ("max_tokens", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("budget_bytes", ctypes.c_uint64),
    ]


class MmcServeStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "step", "admitted", "completed", "block_mallocs", "block_frees", "nulls", "queue_wait_steps", "active",
        "reserved_bytes", "max_active")]


SERVE_SYMBOLS = ["mmc_serve_create", "mmc_serve_destroy", "mmc_serve_load", "mmc_serve_step", "mmc_serve_get_stats",
                 "mmc_serve_timestamps", "mmc_serve_stamp", "mmc_serve_stamp_count"]


def _bind():
    L = mmc.lib()
    if getattr(L, "_serve_bound", False):
        return L
    P = ctypes.POINTER
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    L.mmc_serve_create.argtypes = [vp, P(MmcServeConfig), P(vp)]
    L.mmc_serve_destroy.argtypes = [vp]
    L.mmc_serve_load.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp]
    L.mmc_serve_step.argtypes = [vp, vp]
    L.mmc_serve_get_stats.argtypes = [vp, P(MmcServeStats), vp]
    L.mmc_serve_timestamps.argtypes = [vp, vp, u32, P(ctypes.c_uint64)]
    L.mmc_serve_stamp.argtypes = [vp, vp]
    L.mmc_serve_stamp_count.argtypes = [vp, P(u32), vp]
    for n in SERVE_SYMBOLS:
        getattr(L, n).restype = ctypes.c_int
    L._serve_bound = True
    return L


@dataclasses.dataclass
class Trace:
    arrival: np.ndarray  # uint32, sorted
    length: np.ndarray   # uint32 tokens
    ids: np.ndarray      # uint32 payload ids


def make_trace(seed: int = 5, rate_per_s: float = 2000.0, step_s: float = 0.01, duration_s: float = 30.0,
               mu: float = 6.2, sigma: float = 0.8, max_tokens: int = 8192) -> Trace:
    """Poisson arrivals (rate_per_s) over duration_s in steps of step_s; lengths
    lognormal(mu, sigma) tokens clipped to [1, max_tokens]."""
    rng = np.random.default_rng(seed)
    steps = int(round(duration_s / step_s))
    per_step = rng.poisson(rate_per_s * step_s, size=steps)
    arrival = np.repeat(np.arange(steps, dtype=np.uint32), per_step)
    length = np.clip(np.rint(rng.lognormal(mu, sigma, size=len(arrival))), 1, max_tokens).astype(np.uint32)
    return Trace(arrival, length, np.arange(len(arrival), dtype=np.uint32))


def validate_trace(trace: Trace, max_tokens: int) -> None:
    """The loop's preconditions (include/mallocmc_serve.h, mmc_serve_load):
    every length in 1..max_tokens and arrival steps that never decrease.
    mmc_serve_load checks the same on the device; this names the first
    offending request before anything is copied."""
    length = np.asarray(trace.length)
    arrival = np.asarray(trace.arrival).astype(np.int64)
    if not (len(length) == len(arrival) == len(trace.ids)):
        raise ValueError("trace arrays differ in length")
    bad = np.nonzero((length < 1) | (length > max_tokens))[0]
    if len(bad):
        raise ValueError(f"request {int(bad[0])}: length {int(length[bad[0]])} is outside 1..{max_tokens}")
    down = np.nonzero(np.diff(arrival) < 0)[0]
    if len(down):
        raise ValueError(f"request {int(down[0]) + 1}: arrival steps must not decrease")


def payload_words(rid: int, t: np.ndarray, w: np.ndarray) -> np.ndarray:
    """mmc_serve_word (include/mallocmc_serve.h) over arrays, uint32 arithmetic."""
    with np.errstate(over="ignore"):
        x = (np.uint64(rid) * np.uint64(0x9E3779B1)) & M32
        x = x ^ ((t.astype(np.uint64) * np.uint64(0x85EBCA77)) & M32) ^ ((w.astype(np.uint64) * np.uint64(0xC2B2AE3D)) & M32)
        x = x ^ (x >> np.uint64(15))
        x = (x * np.uint64(0x2C1B3C6D)) & M32
        x = x ^ (x >> np.uint64(12))
    return x


def expected_checksum(rid: int, length: int, token_bytes: int = 256) -> int:
    wpt = token_bytes // 4
    k = np.arange(length * wpt, dtype=np.uint64)
    x = payload_words(rid, k // np.uint64(wpt), k % np.uint64(wpt))
    with np.errstate(over="ignore"):
        return int(np.sum(x * (np.uint64(2) * k + np.uint64(1)), dtype=np.uint64))


def expected_checksums(ids, lengths, token_bytes: int = 256, device: str = "cuda",
                       chunk_terms: int = 1 << 27) -> np.ndarray:
    """expected_checksum of many requests at once, computed with torch on
    `device` in chunks of whole requests.  Every intermediate stays below 2^63
    in int64: a payload word x < 2^32 is split into 16-bit halves, each half
    times (2k + 1) < 2^20 is below 2^36, and a request's at most 2^19 such
    products sum below 2^55; the halves are joined modulo 2^64 at the end."""
    import torch
    wpt = token_bytes // 4
    ids = np.asarray(ids, dtype=np.int64)
    terms = np.asarray(lengths, dtype=np.int64) * wpt
    out = np.zeros(len(ids), dtype=np.uint64)
    M = 0xFFFFFFFF
    start = 0
    while start < len(ids):
        cum = np.cumsum(terms[start:])
        end = start + max(1, int(np.searchsorted(cum, chunk_terms, side="right")))
        n_req = end - start
        seg_terms = torch.from_numpy(terms[start:end]).to(device)
        seg = torch.repeat_interleave(torch.arange(n_req, device=device), seg_terms)
        first = torch.cumsum(seg_terms, 0) - seg_terms
        k = torch.arange(seg.numel(), device=device, dtype=torch.int64) - first[seg]
        rid = torch.from_numpy(ids[start:end]).to(device)[seg]
        x = (rid * 0x9E3779B1) & M
        x = x ^ (((k // wpt) * 0x85EBCA77) & M) ^ (((k % wpt) * 0xC2B2AE3D) & M)
        x = x ^ (x >> 15)
        x = (x * 0x2C1B3C6D) & M
        x = x ^ (x >> 12)
        m = 2 * k + 1
        lo = torch.zeros(n_req, dtype=torch.int64, device=device).index_add_(0, seg, (x & 0xFFFF) * m)
        hi = torch.zeros(n_req, dtype=torch.int64, device=device).index_add_(0, seg, (x >> 16) * m)
        with np.errstate(over="ignore"):
            out[start:end] = lo.cpu().numpy().astype(np.uint64) + (hi.cpu().numpy().astype(np.uint64) << np.uint64(16))
        start = end
    return out


class ServeLoop:
    def __init__(self, allocator: mmc.ScatterAllocator, max_slots: int = 32768, block_bytes: int = 4096,
                 tokens_per_block: int = 16, token_bytes: int = 256, max_tokens: int = 8192,
                 budget_bytes: int | None = None, stamp_capacity: int = 1 << 17):
        import torch
        L = _bind()
        self.alloc = allocator
        g = allocator.geometry()
        if budget_bytes is None:
            budget_bytes = int(g.numpages) * allocator.config.pagesize * 9 // 10
        self.cfg = MmcServeConfig(max_slots, block_bytes, tokens_per_block, token_bytes, max_tokens, 0, budget_bytes)
        self.token_bytes = token_bytes
        self._s = ctypes.c_void_p()
        mmc._check(L.mmc_serve_create(allocator.handle, ctypes.byref(self.cfg), ctypes.byref(self._s)),
                   "mmc_serve_create")
        # device-clock stamp at the start of every step
        self.d_stamps = torch.zeros(stamp_capacity, dtype=torch.int64, device="cuda")
        hz = ctypes.c_uint64()
        mmc._check(L.mmc_serve_timestamps(self._s, ctypes.c_void_p(self.d_stamps.data_ptr()), stamp_capacity,
                                          ctypes.byref(hz)), "mmc_serve_timestamps")
        self.clock_hz = hz.value
        self._stamps = 0

    def close(self):
        if self._s:
            _bind().mmc_serve_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, trace: Trace, stream=None, check: bool = True):
        """Copies the trace to the device and resets the loop; `check`
        validates it on the host first (the library checks it again)."""
        import torch
        if check:
            validate_trace(trace, self.cfg.max_tokens)
        dev = "cuda"
        self.trace = trace
        self.d_arrival = torch.from_numpy(trace.arrival.astype(np.int32)).to(dev)
        self.d_length = torch.from_numpy(trace.length.astype(np.int32)).to(dev)
        self.d_ids = torch.from_numpy(trace.ids.astype(np.int32)).to(dev)
        n = len(trace.arrival)
        self.d_checksums = torch.zeros(max(n, 1), dtype=torch.int64, device=dev)
        self.d_admit = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        mmc._check(_bind().mmc_serve_load(self._s, ctypes.c_void_p(self.d_arrival.data_ptr()),
                                          ctypes.c_void_p(self.d_length.data_ptr()),
                                          ctypes.c_void_p(self.d_ids.data_ptr()), n,
                                          ctypes.c_void_p(self.d_checksums.data_ptr()),
                                          ctypes.c_void_p(self.d_admit.data_ptr()), mmc._stream_ptr(stream)),
                   "mmc_serve_load")
        self._stamps = 0

    def step(self, stream=None):
        mmc._check(_bind().mmc_serve_step(self._s, mmc._stream_ptr(stream)), "mmc_serve_step")

    def stats(self, stream=None) -> dict:
        st = MmcServeStats()
        mmc._check(_bind().mmc_serve_get_stats(self._s, ctypes.byref(st), mmc._stream_ptr(stream)),
                   "mmc_serve_get_stats")
        return {n: getattr(st, n) for n, _ in MmcServeStats._fields_}

    def checksums(self) -> np.ndarray:
        return self.d_checksums.cpu().numpy().view(np.uint64)[: len(self.trace.arrival)]

    def run(self, steps: int, graph: bool = True, time_steps: bool = True, steps_per_graph: int = 100):
        """Runs `steps` steps; returns each step's device time in milliseconds
        (or None).  With `graph`, `steps_per_graph` steps are captured in one
        graph and replayed, so only every steps_per_graph-th step pays a host
        launch.  A step's time runs from its own device-clock stamp to the
        next step's (the last one to a stamp enqueued after it)."""
        import torch
        k = max(1, min(steps_per_graph, steps)) if graph else 1
        n_rep, rest = (steps // k, steps % k) if graph else (0, steps)
        if n_rep:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    for _ in range(k):
                        self.step(stream=s)
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            for _ in range(n_rep):
                g.replay()
        for _ in range(rest):
            self.step()
        mmc._check(_bind().mmc_serve_stamp(self._s, mmc._stream_ptr()), "mmc_serve_stamp")
        i0 = self._stamps
        self._stamps += steps + 1
        torch.cuda.synchronize()
        if not time_steps:
            return None
        if self._stamps > self.d_stamps.numel():
            raise RuntimeError("stamp buffer full: raise stamp_capacity")
        st = self.d_stamps[i0: i0 + steps + 1].cpu().numpy().astype(np.int64)
        return np.diff(st).astype(np.float64) * 1e3 / self.clock_hz


def run_default(log=None, duration_s: float = 30.0, check_sample: int | None = None) -> dict:
    """The default serving workload: 32 GiB heap, 2000 req/s Poisson over 30 s in 10 ms steps,
    lognormal(6.2, 0.8) tokens (<= 8192) at 256 B/token, 4 KiB blocks of 16
    tokens, admission by free bytes.  Runs until every request completed, then
    checks every request's checksum against its closed form (or a seeded
    sample of `check_sample` requests)."""
    import torch
    # one access block over the whole heap: ScatterAlloc fills the current
    # access block before it moves on, and with ~2 GiB live a 2 GiB block
    # turns every block claim into a scan of a nearly full block.  Pages of
    # 64 KiB hold 16 blocks each, so a wave's block claims touch fewer pages:
    # p50 37 us against 45 us at 16 KiB pages and 81 us at 4 KiB
    cfg = mmc.ScatterConfig(pagesize=65536, accessblocksize=32 * GiB, regionsize=32, resetfreedpages=True,
                            flags=mmc.FLAG_FIX_PAGESIZE_LEAK)
    alloc = mmc.ScatterAllocator(cfg, 32 * GiB)
    slots0 = alloc.getAvailableSlots(4096)
    loop = ServeLoop(alloc, max_slots=32768)
    trace = make_trace(duration_s=duration_s)
    loop.load(trace)
    arrival_steps = int(trace.arrival.max()) + 1
    parts = [loop.run(arrival_steps)]
    st = loop.stats()
    while st["completed"] < len(trace.arrival) and len(parts) < 64:
        parts.append(loop.run(1000))  # drain the requests still decoding
        st = loop.stats()
    ms = np.concatenate(parts)
    cs = loop.checksums()
    if check_sample is None:
        sample = np.arange(len(trace.arrival))
    else:
        rng = np.random.default_rng(1)
        sample = rng.choice(len(trace.arrival), size=min(check_sample, len(trace.arrival)), replace=False)
    want = expected_checksums(trace.ids[sample], trace.length[sample], token_bytes=loop.token_bytes)
    bad = int((cs[sample] != want).sum())
    torch.cuda.synchronize()
    slots_ok = alloc.getAvailableSlots(4096) == slots0
    loop.close()
    alloc.close()
    busy = ms[:arrival_steps]  # steps while arrivals continue (steady state)
    total_s = float(ms.sum()) / 1e3
    out = {
        "workload": f"serving loop: Poisson 2000 req/s for {duration_s:.0f} s in 10 ms steps, lognormal(6.2,0.8) tokens "
                    "<= 8192, 256 B/token, 4 KiB blocks, 32 GiB heap (one access block), pagesize 64 KiB, regionsize 32",
        "timing": "device clock stamp at every step start; 100 steps per graph replay (one host launch "
                  "per 100 steps); p50/p99 over the arrival steps",
        "requests": int(len(trace.arrival)), "steps": int(len(ms)), "arrival_steps": arrival_steps,
        "p50_step_us": float(np.percentile(busy, 50) * 1e3), "p99_step_us": float(np.percentile(busy, 99) * 1e3),
        "max_step_us": float(busy.max() * 1e3),
        "steps_per_s": len(ms) / total_s, "allocs_per_s": st["block_mallocs"] / total_s,
        "frees_per_s": st["block_frees"] / total_s,
        "completed": int(st["completed"]), "admitted": int(st["admitted"]), "nulls": int(st["nulls"]),
        "max_active": int(st["max_active"]), "mean_queue_wait_steps": st["queue_wait_steps"] / max(1, st["admitted"]),
        "checksum_mismatches": int(bad), "checksums_checked": int(len(sample)),
        "slots_restored": bool(slots_ok),
    }
    if log:
        log(f"serve: {out}")
    return out
