"""Seeded request workloads of the benchmarked configurations (all integer).

Sizes are drawn with splitmix64 so that every consumer (GPU kernels, tests,
the CPU baseline) reproduces the same request sets from a seed.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finalizer over a uint64 array (Steele, Lea, Flood 2014)."""
    z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def mixed_sizes(seed: int, n: int, steps: int = 256, quantum: int = 16) -> np.ndarray:
    """Uniform sizes {16, 32, ..., 4096} B: 16 * (1 + splitmix64(seed<<32 | tid) mod 256)."""
    tid = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = splitmix64((np.uint64(seed) << np.uint64(32)) | tid)
    return (quantum * (1 + (h % np.uint64(steps)))).astype(np.uint32)


ZIPF_SIZES = np.array([16 << k for k in range(13)], dtype=np.uint32)  # 16 .. 65536


def zipf_sizes(rng: np.random.Generator, n: int, s: float = 1.2) -> np.ndarray:
    """Zipf(s) over the powers of two 16..65536 B."""
    w = 1.0 / np.arange(1, len(ZIPF_SIZES) + 1) ** s
    p = w / w.sum()
    return ZIPF_SIZES[rng.choice(len(ZIPF_SIZES), size=n, p=p)]


def bimodal_sizes(rng: np.random.Generator, n: int) -> np.ndarray:
    """80% uniform 16-256 B, 20% uniform 8-60 KiB, padded to 16 B."""
    small = rng.integers(16, 257, size=n)
    large = rng.integers(8 * 1024, 60 * 1024 + 1, size=n)
    pick = rng.random(n) < 0.8
    s = np.where(pick, small, large)
    return (((s + 15) // 16) * 16).astype(np.uint32)


def random_stream(rng: np.random.Generator, n: int, sizes, p_free: float = 0.35) -> np.ndarray:
    """Serialized op stream: op >= 0 is malloc(op); op < 0 frees request (-op-1).
    Every malloc is freed at most once and only after it was issued."""
    ops = []
    live = []
    sizes = list(sizes)
    si = 0
    while len(ops) < n:
        if live and rng.random() < p_free:
            j = int(rng.integers(len(live)))
            live[j], live[-1] = live[-1], live[j]
            ops.append(-(live.pop()) - 1)
        else:
            ops.append(int(sizes[si % len(sizes)]))
            live.append(len(ops) - 1)
            si += 1
    return np.array(ops, dtype=np.int64)
