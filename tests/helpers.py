"""Test helpers: the reference's mulberry32 stream (test/gpu/renderUtils.ts:23-30)
vectorised (its state only advances by a constant), and comparisons."""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint32(0x6D2B79F5)


class Rng:
    """mulberry32(seed): .random(n) gives the next n numbers in [0, 1)."""

    def __init__(self, seed: int):
        self.state = np.uint32(seed & 0xFFFFFFFF)

    def random(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            s = (np.uint64(self.state) + k * np.uint64(GOLDEN)) & np.uint64(0xFFFFFFFF)
            s = s.astype(np.uint32)
            t = (s ^ (s >> np.uint32(15))) * (np.uint32(1) | s)
            t = (t + ((t ^ (t >> np.uint32(7))) * (np.uint32(61) | t))) ^ t
            out = (t ^ (t >> np.uint32(14))).astype(np.float64) / 4294967296.0
            self.state = np.uint32((int(self.state) + n * int(GOLDEN)) & 0xFFFFFFFF)
        return out


def first_mismatch(actual, expected):
    a = np.asarray(actual)
    e = np.asarray(expected)
    if a.shape != e.shape:
        return f"shape {a.shape} vs {e.shape}"
    bad = np.nonzero(a != e)[0]
    if bad.size:
        i = int(bad[0])
        return f"index {i}: got {a[i]}, expected {e[i]} ({bad.size} differ)"
    return None
