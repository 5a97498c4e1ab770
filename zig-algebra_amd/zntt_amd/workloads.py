"""Seeded synthetic workloads of the benchmark configurations.

Metadata (sizes, directions, arrival times) comes from a xoshiro256++ stream
seeded by SplitMix64 -- the published generator, restated here in a few
lines of Python because it draws at most a few tens of thousands of values.
Coefficients of item i come from the library's own generator
(``zntt_fill_uniform``) with the per-item seed ``(seed << 32) | (i + 1)``, so
items can be generated independently and in parallel.
"""
from __future__ import annotations

import math

_M64 = (1 << 64) - 1


class Xoshiro256pp:
    def __init__(self, seed: int):
        st = seed & _M64
        self.s = []
        for _ in range(4):
            st = (st + 0x9E3779B97F4A7C15) & _M64
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            self.s.append(z ^ (z >> 31))

    def next(self) -> int:
        s = self.s
        r = ((((s[0] + s[3]) & _M64) << 23 | ((s[0] + s[3]) & _M64) >> 41) & _M64)
        r = (r + s[0]) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & _M64
        return r

    def uniform_int(self, lo: int, hi: int) -> int:
        """Uniform in [lo, hi] by rejection (no modulo bias)."""
        span = hi - lo + 1
        limit = (1 << 64) - ((1 << 64) % span)
        while True:
            x = self.next()
            if x < limit:
                return lo + x % span

    def uniform01(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))


def item_seed(seed: int, i: int) -> int:
    return (seed << 32) | (i + 1)


def uneven_batch(count: int = 512, lo: int = 10, hi: int = 20, seed: int = 3) -> list[tuple[int, int]]:
    """(log_n, direction) per transform; direction 0 forward, 1 inverse,
    2 coset-forward, each with probability 1/3; all metadata drawn first."""
    g = Xoshiro256pp(seed)
    sizes = [g.uniform_int(lo, hi) for _ in range(count)]
    dirs = [g.uniform_int(0, 2) for _ in range(count)]
    return list(zip(sizes, dirs))


def serve_arrivals(rate: float = 2000.0, duration: float = 10.0, k_lo: int = 12, k_hi: int = 22,
                   seed: int = 4) -> list[tuple[float, int, int]]:
    """(arrival time in s, log2 coefficient count k, log2 blowup b) per request:
    Poisson arrivals at `rate`, k uniform in [k_lo, k_hi], blowup 4 or 8."""
    g = Xoshiro256pp(seed)
    out = []
    t = 0.0
    while True:
        t += -math.log(1.0 - g.uniform01()) / rate
        if t >= duration:
            break
        k = g.uniform_int(k_lo, k_hi)
        b = 2 + g.uniform_int(0, 1)
        out.append((t, k, b))
    return out
