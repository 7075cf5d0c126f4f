"""The xgminer front end and the edge-list loader on the CPU: option
handling and error exits of the command line (no device is touched before
the graph loads), and the edge-list text format (xgminer/src/dataloader.cpp:
29-175) against the reference's citeseer graph and its known counts."""
import os
import subprocess

import numpy as np
import pytest

from conftest import CITESEER


def _run(xgm, *args):
    return subprocess.run([xgm.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def _write_edgelist(path, n, edges):
    with open(path, "w") as f:
        f.write(f"{n} {len(edges)}\n")
        for u, v in edges:
            f.write(f"{u} {v}\n")


def test_cli_options(xgm):
    assert os.path.exists(xgm.CLI_PATH)
    r = _run(xgm, "--help")
    assert r.returncode == 0 and "--patternID" in r.stdout
    for args, msg in ((("-g", CITESEER + "_missing", "--patternID", "20"), "cannot load graph"),
                      (("-g", os.path.dirname(CITESEER), "--patternID", "22"), "unsupported --patternID"),
                      (("-g", os.path.dirname(CITESEER), "--bogus", "1"), "unknown option"),
                      (("-g", os.path.dirname(CITESEER), "--algorithm", "cpu_baseline", "--patternID", "20"),
                       "no CPU path"),
                      (("-g", os.path.dirname(CITESEER), "--pattern-size", "4", "--pattern-adj-mat",
                        "0110101111011010"), "unsupported pattern matrix"),
                      (("--patternID", "20"), "no graph given")):
        r = _run(xgm, *args)
        assert r.returncode == 1 and msg in r.stderr, (args, r.stderr)


def test_edgelist_first_appearance_ids(xgm, tmp_path):
    p = str(tmp_path / "g.txt")
    with open(p, "w") as f:
        f.write("3 4\n10 20\n20 5\n7 7\n5 20\n")  # a self-loop line and a repeated edge
    g = xgm.load_edgelist(p)
    assert list(g.rowptr) == [0, 1, 3, 4]
    assert list(g.colidx) == [1, 0, 2, 1]


def test_edgelist_header_checks(xgm, tmp_path):
    p = str(tmp_path / "g.txt")
    for text in ("3 3\n1 2\n2 3\n", "4 2\n1 2\n2 3\n", "x y\n"):
        with open(p, "w") as f:
            f.write(text)
        with pytest.raises(xgm.XGMError):
            xgm.load_edgelist(p)


def test_edgelist_citeseer(xgm, orc, tmp_path):
    g = xgm.load_csr(CITESEER)
    deg = np.diff(g.rowptr)
    src = np.repeat(np.arange(g.n), deg)
    keep = src < g.colidx
    rng = np.random.default_rng(4)
    ids = rng.permutation(10 ** 6)[: g.n] + 5000  # arbitrary sparse ids
    edges = list(zip(ids[src[keep]], ids[g.colidx[keep]]))
    rng.shuffle(edges)
    n_used = int((deg > 0).sum())
    p = str(tmp_path / "citeseer.txt")
    _write_edgelist(p, n_used, edges)
    h = xgm.load_edgelist(p)
    assert h.n == n_used and h.nnz == g.nnz
    assert sorted(np.diff(h.rowptr)) == sorted(deg[deg > 0])
    d = h.orient()
    assert orc.count("triangle", d.rowptr, d.colidx)[0] == 1166
    assert orc.count("clique4", d.rowptr, d.colidx)[0] == 255
    assert orc.count("diamond", h.rowptr, h.colidx)[0] == 3730
    assert orc.count("rectangle", h.rowptr, h.colidx)[0] == 6059
