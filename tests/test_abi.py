"""The C-ABI library loads and exports every function the public headers
declare; CPU-side error paths return status codes (no GPU compute here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared():
    names = []
    for h in ("xgm.h", "xgm_graph.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(xgm_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_exports_every_declared_symbol(xgm):
    names = _declared()
    assert "xgm_count" in names and "xgm_gen_rmat" in names
    lib = ctypes.CDLL(xgm.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror(xgm):
    assert xgm.lib.xgm_strerror(0) == b"ok"
    assert xgm.lib.xgm_strerror(3) == b"unsupported"


def test_null_arguments_rejected(xgm):
    assert xgm.lib.xgm_count(None, None, 1, None, None) == xgm.XGM_EINVAL
    assert xgm.lib.xgm_graph_upload(None, None, None) == xgm.XGM_EINVAL
    assert xgm.lib.xgm_graph_orient(None, None, None) == xgm.XGM_EINVAL
    assert xgm.lib.xgm_graph_download(None, None, None) == xgm.XGM_EINVAL
    assert xgm.lib.xgm_graph_task_order(None, None, 0, None) == xgm.XGM_EINVAL
    assert xgm.lib.xgm_gen_rmat(0, 16, 0.57, 0.19, 0.19, 1, None) == xgm.XGM_EINVAL
    with pytest.raises(xgm.XGMError):
        xgm.gen_rmat(12, 16, 0.9, 0.9, 0.9, 1)  # probabilities sum > 1


def test_solver_names_mirror_reference(xgm):
    for f in ("TCSolver", "CliqueSolver", "SglSolver"):
        assert callable(getattr(xgm, f))
    with pytest.raises(xgm.XGMError):
        xgm.CliqueSolver(None, 7)
    with pytest.raises(xgm.XGMError):
        xgm.SglSolver(None, "house")
