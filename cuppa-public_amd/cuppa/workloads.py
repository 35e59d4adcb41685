"""The synthetic workloads of the benchmark, shared by bench.py and the tests.

* ``batch_specs``: the batched ensemble -- 4096 independent copy-model graphs,
  n log-uniform in [1e3, 1e7], d uniform in 1..16, p uniform in [0.05, 0.95],
  Philox seed 1000 + i.
* ``serve_trace``: the request mix of the serving loop -- n lognormal (median
  1e5, sigma 1.5) clipped to [1e2, 1e8], d in {2, 4, 8, 16}, p in {0.1, 0.5,
  0.9}, seed = request id.
* ``poisson_arrivals``: arrival times (ms) of a Poisson process at a rate.

The meta-draws come from numpy's Philox generator with a fixed key, so every
run sees the same graphs. Nothing here computes graph values.
"""
from __future__ import annotations

import numpy as np


def _p_threshold(p: float) -> int:
    """min(2^32, floor(p * 2^32)), as ppa_p_threshold computes it."""
    if not p > 0.0:
        return 0
    if p >= 1.0:
        return 1 << 32
    return int(p * 4294967296.0)


def batch_specs(count: int = 4096):
    """[(n, d, p_threshold, seed)] of the ensemble; the meta-draw is numpy's
    Philox generator keyed 4096."""
    rng = np.random.Generator(np.random.Philox(key=4096))
    out = []
    for i in range(count):
        n = int(round(10 ** rng.uniform(3, 7)))
        d = int(rng.integers(1, 17))
        p = float(rng.uniform(0.05, 0.95))
        out.append((max(n, d + 1), d, _p_threshold(p), 1000 + i))
    return out


def serve_trace(count: int, seed: int = 5):
    """[(id, n, d, p_threshold, seed)] of the request mix; request i's Philox
    seed is i."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    n = np.clip(np.rint(rng.lognormal(np.log(1e5), 1.5, count)), 1e2, 1e8).astype(np.int64)
    d = rng.choice([2, 4, 8, 16], count)
    p = rng.choice([0.1, 0.5, 0.9], count)
    thr = {q: _p_threshold(q) for q in (0.1, 0.5, 0.9)}
    return [(i, int(max(n[i], d[i] + 1)), int(d[i]), thr[float(p[i])], i) for i in range(count)]


def poisson_arrivals(count: int, rate_per_s: float, key: int = 1):
    """Arrival times in ms of `count` requests at `rate_per_s` (0: all at 0)."""
    if not rate_per_s:
        return [0.0] * count
    rng = np.random.Generator(np.random.Philox(key=key))
    return np.cumsum(rng.exponential(1e3 / rate_per_s, count)).tolist()
