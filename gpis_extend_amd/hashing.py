"""Vectorised host restatement of the reference's seed hashes and PCG32 sampler
(src/core/math/MathUtil.hpp:179-224, src/core/sampling/UniformSampler.hpp:17-75,
src/core/math/BitManip.hpp:47-50), used to build flight requests on the host."""
from __future__ import annotations

import numpy as np

P2, P3, P4, P5 = np.uint32(2246822519), np.uint32(3266489917), np.uint32(668265263), np.uint32(374761393)
_MUL = np.uint64(6364136223846793005)


def _u32(a):
    return np.asarray(a).astype(np.uint32)


def _rotl17(h):
    return (h << np.uint32(17)) | (h >> np.uint32(15))


def _final(h):
    h = P2 * (h ^ (h >> np.uint32(15)))
    h = P3 * (h ^ (h >> np.uint32(13)))
    return h ^ (h >> np.uint32(16))


def xxhash32_1(p):
    with np.errstate(over="ignore"):
        return _final(P4 * _rotl17(_u32(p) + P5))


def xxhash32_2(x, y):
    with np.errstate(over="ignore"):
        return _final(P4 * _rotl17(_u32(y) + P5 + _u32(x) * P3))


def xxhash32_4(x, y, z, w):
    with np.errstate(over="ignore"):
        h = P4 * _rotl17(_u32(w) + P5 + _u32(x) * P3)
        h = P4 * _rotl17(h + _u32(y) * P3)
        h = P4 * _rotl17(h + _u32(z) * P3)
        return _final(h)


class Pcg32:
    """UniformSampler with set_state semantics (two discarded draws), vectorised over lanes."""

    def __init__(self, state):
        self.s = np.asarray(state).astype(np.uint64).copy()
        self.next_u32()
        self.next_u32()

    def next_u32(self):
        with np.errstate(over="ignore"):
            old = self.s
            self.s = old * _MUL + np.uint64(1)
            xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
            rot = (old >> np.uint64(59)).astype(np.uint32)
            return (xs >> rot) | (xs << ((np.uint32(0) - rot) & np.uint32(31)))

    def next_1d(self):
        u = (self.next_u32() >> np.uint32(9)) | np.uint32(0x3F800000)
        return u.view(np.float32) - np.float32(1.0)
