"""Mixed-size workloads for the batched and served modes.

- Batched: independent directed DCSBM graphs.  Graph i draws N log-uniform in [200, 50,000],
  B = round(N^0.4) equal blocks, a mean degree uniform in [8, 32] (power law, exponent 2.5, on
  [1, min(N / 2, 1000)]) and an intra-block fraction uniform in [0.5, 0.8]; it is generated with seed i.
- Served: requests arrive as a Poisson process (arrival seed 0).  Request j draws N log-uniform in
  [1,000, 2,000,000] (or a smaller upper bound), B = round(N^0.4) equal blocks, mean degree 16 (power law,
  exponent 2.1, on [1, min(N / 2, 2000)]), intra-block fraction 0.55; it is generated with seed j.

The size and shape draws of graph i / request j come from their own Philox stream (key 2^32 + i, resp.
2^33 + j), apart from the generator's.
"""
import concurrent.futures as cf
import math
import os

import numpy as np

from .graph import dcsbm


def _draws(key):
    return np.random.Generator(np.random.Philox(key))


def batched_spec(i, n_min=200, n_max=50_000):
    r = _draws((1 << 32) + i)
    N = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
    return dict(N=N, B=max(1, int(round(N ** 0.4))), gamma=2.5, kmin=1, kmax=max(2, min(N // 2, 1000)),
                mean_degree=float(r.uniform(8.0, 32.0)), p_in=float(r.uniform(0.5, 0.8)), seed=i)


def served_spec(j, n_min=1_000, n_max=2_000_000):
    r = _draws((1 << 33) + j)
    N = int(round(math.exp(r.uniform(math.log(n_min), math.log(n_max)))))
    return dict(N=N, B=max(1, int(round(N ** 0.4))), gamma=2.1, kmin=1, kmax=max(2, min(N // 2, 2000)),
                mean_degree=16.0, p_in=0.55, seed=j)


def make(spec):
    s = dict(spec)
    N, B = s.pop("N"), s.pop("B")
    return dcsbm(N, B, **s)


def generate(specs, workers=None):
    """The graphs of `specs`, generated in parallel processes."""
    workers = workers or min(8, os.cpu_count() or 1, max(1, len(specs)))
    if workers <= 1 or len(specs) <= 1:
        return [make(s) for s in specs]
    with cf.ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(make, specs, chunksize=1))


def poisson_arrivals(rate, count, seed=0):
    """Arrival times in seconds of `count` requests at `rate` per second (exponential gaps, first at 0)."""
    gaps = np.random.Generator(np.random.Philox(seed)).exponential(1.0 / rate, count)
    t = np.cumsum(gaps)
    return t - t[0]
