"""Seeded synthetic inputs of the QGT workloads (no GPU, no library calls).

All draws come from one splitmix64 stream per workload: x -> (x >> 11) 2^-53
in [0, 1); an angle is 2 pi u; an integer in {lo..hi} is lo + floor(u (hi - lo + 1)).
"""
import math

import numpy as np

_M = 0xFFFFFFFFFFFFFFFF


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _M

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def integer(self, lo, hi):
        return lo + min(hi - lo, int(self.uniform() * (hi - lo + 1)))

    def angles(self, count):
        return np.array([2.0 * math.pi * self.uniform() for _ in range(count)], np.float64)


def hea_batch(count, seed=0, n_range=(12, 18), layer_range=(2, 8)):
    """Batched HEA workload: per instance, in order, n ~ U{n_range}, L ~ U{layer_range},
    then the 2 n L angles, all from one stream.  Returns [(n, L, theta), ...]."""
    g = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = g.integer(*n_range)
        L = g.integer(*layer_range)
        out.append((n, L, g.angles(2 * n * L)))
    return out
