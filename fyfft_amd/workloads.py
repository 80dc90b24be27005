"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

ragged_lengths: the grouped ragged configuration -- `count` lengths, a class
mix of 60% 2/3/5/7-smooth lengths, 25% Rader primes (p - 1 smooth) and 15%
Bluestein primes (p - 1 not smooth), log-uniformly spread from 16 upward.

Rule for the total size.  The configuration asks for lengths log-uniform on
[16, 2^20] AND a total of about 2^28 elements for 20000 transforms; a
log-uniform draw on [16, 2^20] has mean (2^20 - 16) / ln(2^16) ~ 94.5K, so
20000 draws sum to ~1.9e9, not 2^28.  Here the log-uniform variate is mapped
onto [16, L_max] with L_max solved so that the expected total is
`total_target` (L_max ~ 1.2e5 for 20000 transforms and 2^28), i.e.
n = 16 * (L_max / 16)^u with u ~ U[0, 1); `L_max = 2**20` reproduces the pure
draw.  Each draw is then snapped to the nearest length of its class (in log
distance).
"""
from __future__ import annotations

import bisect
import math

import numpy as np


def _is_smooth(n: int) -> bool:
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def _tables(limit: int):
    sieve = np.ones(limit + 1, bool)
    sieve[:2] = False
    for i in range(2, int(limit ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    primes = np.nonzero(sieve)[0]
    smooth = [n for n in range(16, limit + 1) if _is_smooth(n)]
    rader = [int(p) for p in primes if p >= 16 and _is_smooth(int(p) - 1)]
    blue = [int(p) for p in primes if p >= 16 and not _is_smooth(int(p) - 1)]
    return smooth, rader, blue


def _nearest(sorted_vals, x: float) -> int:
    i = bisect.bisect_left(sorted_vals, x)
    cands = [sorted_vals[j] for j in (i - 1, i) if 0 <= j < len(sorted_vals)]
    return min(cands, key=lambda v: abs(math.log(v / x)))


def l_max_for(count: int, total_target: float, lo: float = 16.0) -> float:
    """L_max with count * E[log-uniform on [lo, L_max]] = total_target."""
    target = total_target / count
    a, b = lo * 1.01, float(1 << 30)
    for _ in range(200):
        m = math.sqrt(a * b)
        mean = (m - lo) / math.log(m / lo)
        if mean < target:
            a = m
        else:
            b = m
    return math.sqrt(a * b)


def ragged_lengths(count: int = 20000, seed: int = 0, total_target: float = 2.0 ** 28,
                   l_max: float | None = None, mix=(0.60, 0.25, 0.15)) -> np.ndarray:
    if l_max is None:
        l_max = l_max_for(count, total_target)
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, count)
    raw = 16.0 * (l_max / 16.0) ** u
    cls = rng.choice(3, size=count, p=list(mix))
    smooth, rader, blue = _tables(int(l_max * 1.5) + 64)
    out = np.empty(count, np.int64)
    for i in range(count):
        table = (smooth, rader, blue)[cls[i]]
        out[i] = _nearest(table, raw[i])
    return out
