"""Shared test fixtures.

The package directory is ``x-gminer_amd/`` (not an importable identifier), so
it is loaded by path as module ``xgminer_amd``.  The oracle package lives in
``oracle/`` at the repository root."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
CITESEER = os.path.join(GOLDEN, "citeseer", "graph")


def load_xgm():
    if "xgminer_amd" in sys.modules:
        return sys.modules["xgminer_amd"]
    pkg = os.path.join(ROOT, "x-gminer_amd")
    spec = importlib.util.spec_from_file_location(
        "xgminer_amd", os.path.join(pkg, "__init__.py"), submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["xgminer_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")


@pytest.fixture(scope="session")
def xgm():
    return load_xgm()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def complete_graph(xgm, n):
    src, dst = np.triu_indices(n, 1)
    return xgm.from_edges(n, src, dst)


def complete_bipartite(xgm, a, b):
    src = np.repeat(np.arange(a), b)
    dst = a + np.tile(np.arange(b), a)
    return xgm.from_edges(a + b, src, dst)


def random_graph(xgm, n, m, seed):
    rng = np.random.default_rng(seed)
    return xgm.from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m))
