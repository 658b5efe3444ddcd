"""Seeded operating points (Mach number, angle of attack) for batched sweeps and the learned
warm start's training and held-out sets.

Draws follow C++ `std::mt19937_64` with `std::uniform_real_distribution<double>(a, b)` as
libstdc++ implements it (one 64-bit draw per value: a + (b - a) * (x / 2^64)). Each point takes
two consecutive draws, the Mach number first, then the angle of attack in degrees.
"""
from __future__ import annotations

import numpy as np

_MASK = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne twister (Matsumoto and Nishimura), parameters of std::mt19937_64."""

    N, M = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF

    def __init__(self, seed: int):
        mt = [0] * self.N
        mt[0] = seed & _MASK
        for i in range(1, self.N):
            mt[i] = (6364136223846793005 * (mt[i - 1] ^ (mt[i - 1] >> 62)) + i) & _MASK
        self.mt, self.i = mt, self.N

    def _twist(self):
        mt = self.mt
        for k in range(self.N):
            y = (mt[k] & self.UPPER) | (mt[(k + 1) % self.N] & self.LOWER)
            mt[k] = mt[(k + self.M) % self.N] ^ (y >> 1) ^ (self.MATRIX_A if y & 1 else 0)
        self.i = 0

    def __call__(self) -> int:
        if self.i >= self.N:
            self._twist()
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & _MASK

    def uniform(self, a: float, b: float) -> float:
        u = float(self()) / 18446744073709551616.0
        if u >= 1.0:
            u = np.nextafter(1.0, 0.0)
        return a + (b - a) * u


def operating_points(seed: int, count: int, mach=(0.3, 0.75), aoa_deg=(-4.0, 6.0)):
    """`count` points (Mach, AoA in degrees) from std::mt19937_64(seed)."""
    g = MT19937_64(seed)
    m = np.empty(count)
    a = np.empty(count)
    for i in range(count):
        m[i] = g.uniform(*mach)
        a[i] = g.uniform(*aoa_deg)
    return m, a
