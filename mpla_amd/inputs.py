"""Host-side synthetic inputs with the library's counter-based generator.

SplitMix64 (Steele, Lea & Flood 2014) keyed by (seed, stream): output i is
mix64(key + (i + 1) * GOLDEN) with key = mix64(seed * GOLDEN + stream), mapped
to [0, 1) with 53 bits.  mplak_fill_uniform_f64 produces the same doubles on
the device.  Normals use Box-Muller on outputs 2i and 2i + 1.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _key(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        return _mix64(np.uint64(seed) * _GOLDEN + np.uint64(stream))


def u64(seed: int, stream: int, first: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        i = np.arange(first, first + n, dtype=np.uint64)
        return _mix64(_key(seed, stream) + (i + np.uint64(1)) * _GOLDEN)


def uniform(n: int, seed: int, stream: int, lo: float = 0.0, hi: float = 1.0, first: int = 0) -> np.ndarray:
    """lo + (hi - lo) * U; bitwise equal to the device generator whenever
    (hi - lo) * U is exact (spans that are powers of two, e.g. [0,1], [-1,1])."""
    u = (u64(seed, stream, first, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return lo + (hi - lo) * u


def normal(n: int, seed: int, stream: int, first: int = 0) -> np.ndarray:
    w = u64(seed, stream, 2 * first, 2 * n)
    u1 = ((w[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (w[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
