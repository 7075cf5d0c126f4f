"""Synthetic workloads: the graphs the counts and timings are quoted on.

* ``rmat(scale, seed)``: RMAT, edge factor 16, (a, b, c, d) = (0.57, 0.19,
  0.19, 0.05), symmetrised, deduplicated, self-loops removed.
* ``chung_lu_4m()``: Chung-Lu power law, n = 4,000,000, average degree 20,
  exponent 2.1, expected degree capped at 50,000, seed 3.
* ``batch_specs()`` / ``batch_graphs()``: 256 instances of uneven size,
  seeds 100..355.  Seeds 100..227 are Erdos-Renyi with n = round(10^U(3,
  log10 2e5)) and average degree U[4, 64]; seeds 228..355 are RMAT with scale
  U{10..18} and edge factor 16.  The draws come from numpy's PCG64 seeded with
  the instance seed, so the list is the same on every machine.
"""
import math

import numpy as np

RMAT_ABC = (0.57, 0.19, 0.19)
BATCH_SEEDS = range(100, 356)


def rmat(xgm, scale, seed, edge_factor=16):
    return xgm.gen_rmat(scale, edge_factor, *RMAT_ABC, seed)


def chung_lu_4m(xgm):
    return xgm.gen_chung_lu(4_000_000, 20.0, 2.1, 50_000.0, 3)


def batch_specs(count=None):
    """[(seed, kind, params)] for the 256 batch instances (first `count`
    of each half when count is given)."""
    specs = []
    for seed in BATCH_SEEDS:
        rng = np.random.default_rng(seed)
        if seed < 228:
            n = int(round(10 ** rng.uniform(3.0, math.log10(2e5))))
            d = float(rng.uniform(4.0, 64.0))
            specs.append((seed, "er", dict(n=n, avg_degree=d)))
        else:
            specs.append((seed, "rmat", dict(scale=int(rng.integers(10, 19)))))
    if count is not None:
        er = [s for s in specs if s[1] == "er"][:count]
        rm = [s for s in specs if s[1] == "rmat"][:count]
        specs = er + rm
    return specs


def make_instance(xgm, spec):
    seed, kind, p = spec
    if kind == "er":
        return xgm.gen_erdos_renyi(p["n"], p["avg_degree"], seed)
    return rmat(xgm, p["scale"], seed)


def batch_graphs(xgm, specs=None):
    """Undirected instances of the batch, in spec order."""
    return [make_instance(xgm, s) for s in (specs or batch_specs())]
