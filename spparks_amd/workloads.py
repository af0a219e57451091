"""Synthetic multi-lattice workloads: the 512-replica ensemble and the
Poisson stream of grain-growth solves.  Pure functions of a fixed numpy
generator, so every run (and every test) sees the same replicas."""
from __future__ import annotations

import math

import numpy as np

ENSEMBLE_SEED = 20261202


def even_edge(x):
    """2 * round(x / 2): the nearest even lattice edge."""
    return int(2 * round(x / 2.0))


def ensemble_replicas(n_ising=256, n_potts=256, sweeps=100, seed=ENSEMBLE_SEED):
    """The batched-ensemble workload: n_ising 2d Ising replicas on sq/4n with
    edges log-uniform in 64..1024 and kT uniform in 1.5..3.5, then n_potts 3d
    Potts replicas on sc/26n with edges log-uniform in 32..128, Q log-uniform
    in 100..1e5 and kT uniform in 0.2..1.0.  Replica r has master seed
    1000 + r; kT is rounded to 4 decimals as a script would carry it."""
    rng = np.random.default_rng(seed)
    li = rng.uniform(math.log(64), math.log(1024), n_ising)
    ti = rng.uniform(1.5, 3.5, n_ising)
    lp = rng.uniform(math.log(32), math.log(128), n_potts)
    qp = rng.uniform(math.log(100), math.log(1e5), n_potts)
    tp = rng.uniform(0.2, 1.0, n_potts)
    reps = []
    for r in range(n_ising):
        L = even_edge(math.exp(li[r]))
        reps.append({"id": r, "app": "ising", "lattice": "sq/4n", "shape": [L, L, 1], "nspins": 2,
                     "kT": round(float(ti[r]), 4), "seed": 1000 + r, "sweeps": sweeps})
    for k in range(n_potts):
        r = n_ising + k
        L = even_edge(math.exp(lp[k]))
        reps.append({"id": r, "app": "potts", "lattice": "sc/26n", "shape": [L, L, L],
                     "nspins": int(round(math.exp(qp[k]))), "kT": round(float(tp[k]), 4),
                     "seed": 1000 + r, "sweeps": sweeps})
    return reps


def serving_requests(rate, n, seed=ENSEMBLE_SEED + 5):
    """A Poisson stream of n 3d Potts grain-growth requests at `rate`/s on
    sc/26n: edges even-uniform in 32..192, Q log-uniform in 100..1e6, kT
    uniform in 0.1..1.0, sweeps uniform in 50..500, seed 10^6 + request id.
    Returns requests sorted by arrival time (seconds from 0)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arrive = np.cumsum(gaps) - gaps[0]
    edges = 2 * rng.integers(16, 97, n)            # 32..192, even
    q = np.rint(np.exp(rng.uniform(math.log(100), math.log(1e6), n))).astype(int)
    kt = rng.uniform(0.1, 1.0, n)
    sw = rng.integers(50, 501, n)
    return [{"id": i, "arrive": float(arrive[i]), "app": "potts", "lattice": "sc/26n",
             "shape": [int(edges[i])] * 3, "nspins": int(q[i]), "kT": round(float(kt[i]), 4),
             "seed": 10**6 + i, "sweeps": int(sw[i])} for i in range(n)]
