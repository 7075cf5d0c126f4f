"""Learned clique scheduler: per-task cost model over the strategy codes of
``xgm_sched_create`` (0 id-ordered LUT, 1 rank-ordered LUT, 2 edge LUT,
3 edge nested lists, 4 rank-ordered LUT one capacity bin up).

Codes 1 and 4 on vertices with 1024 < d+ <= 4096 run the rank-ordered LUT
for hubs (rows packed upper-triangular in LDS up to 1344 out-neighbours, in
HBM above); the plain heuristic sends those vertices to the edge LUT.

* ``collect``: per-vertex features (computed on the GPU) and per-vertex
  cost under every code, from instrumented runs of the plain kernels: the
  task's wall time divided by the resident blocks of its kernel, i.e. its
  share of the kernel duration.  Targets are log2(1 + ns); rows a code cannot run
  (LUT codes on vertices above 1024 out-neighbours) are masked.
* ``train``: fits the 16-64-64-5 MLP on the GPU (the library's own BF16-MFMA
  forward/backward/Adam kernels).
* ``schedule``: features, batched inference and argmin on the GPU -> codes.

The model never changes a count: every code counts every task exactly once.
The degree-threshold heuristic (``DeviceGraph.sched_heuristic``) is the
fallback whenever no model is given.  The reference has no learned
scheduler; its kernels use fixed thresholds (xgminer/include/common.h:61,
114-116).
"""
from __future__ import annotations

import os

import numpy as np

from . import CLIQUE5, MLP, Schedule

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "k5_cost_mlp.npy")
NCODES = 5


def collect(dg, pattern=CLIQUE5, codes=range(NCODES), max_rows=None, seed=0):
    """Features X (m x 16), targets T and masks M (m x 5) for the task
    vertices of one device graph, and the exact count (equal for every code)."""
    F = dg.sched_features(pattern)
    task = F[:, 15] > 0
    dplus = np.diff(dg.host.rowptr)
    T = np.zeros((dg.n, NCODES), np.float32)
    Mk = np.zeros((dg.n, NCODES), np.float32)
    count = None
    for c in codes:
        cv = np.where(task, c, -1).astype(np.int8)
        s = Schedule(dg, pattern, cv)
        n, us = s.count(log=True)
        s.close()
        if count is None:
            count = n
        elif n != count:
            raise RuntimeError(f"count differs between schedules: {n} vs {count} (code {c})")
        T[:, c] = np.log2(1.0 + 1000.0 * us)  # kernel-time share in ns
        limit = 1024 if c == 0 else 4096  # id-ordered LUT in LDS; rank LUT up to 4096 rows (HBM above 1024)
        valid = task & ((c in (2, 3)) | (dplus <= limit))
        Mk[:, c] = valid
    rows = np.nonzero(task)[0]
    if max_rows and len(rows) > max_rows:
        rng = np.random.default_rng(seed)
        # keep every heavy task, sample the light ones
        w = F[rows, 1]
        heavy = rows[w >= np.quantile(w, 0.9)]
        light = rows[w < np.quantile(w, 0.9)]
        take = max(0, max_rows - len(heavy))
        rows = np.concatenate([heavy, rng.choice(light, size=min(take, len(light)), replace=False)])
    return F[rows], T[rows], Mk[rows], count


def train(X, T, M, epochs=40, batch=4096, lr=2e-3, seed=1, device=0):
    model = MLP(device, seed=seed)
    loss = model.train(X, T, M, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return model, loss


def load(path=DEFAULT_MODEL, device=0):
    model = MLP(device, seed=1)
    model.params = np.load(path, allow_pickle=False)
    return model


def save(model, path=DEFAULT_MODEL):
    np.save(path, model.params)


def schedule(dg, model, pattern=CLIQUE5):
    """Learned codes (GPU features + inference + argmin); None -> heuristic."""
    if model is None:
        return dg.sched_heuristic(pattern)
    return dg.sched_learned(pattern, model)


def oracle_best(T, M):
    """Per-row best code under the measured costs (upper bound for a model)."""
    Tm = np.where(M > 0, T, np.inf)
    return np.argmin(Tm, axis=1)
