"""The xgminer front end on the GPU: the reference's citeseer graph in its
binary CSR directory and as edge-list text, every supported way of naming a
pattern, against the reference's known counts (codes/GraphMiner READMEs,
src/motif/README.md:53-66; 5-cliques 46 by brute force)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import CITESEER

pytestmark = pytest.mark.gpu

K4 = "0111101111011110"
DIAMOND = "0111101111001100"
CYCLE4 = "0101101001011010"


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _counts(xgm, *args):
    r = subprocess.run([xgm.CLI_PATH, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (args, r.stdout, r.stderr)
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"Pattern (\S+) count: (\d+)", r.stdout)}


def _edgelist(xgm, tmp_path):
    g = xgm.load_csr(CITESEER)
    src = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    keep = src < g.colidx
    p = str(tmp_path / "citeseer.txt")
    with open(p, "w") as f:
        f.write(f"{int((np.diff(g.rowptr) > 0).sum())} {int(keep.sum())}\n")
        for u, v in zip(src[keep], g.colidx[keep]):
            f.write(f"{u + 100} {v + 100}\n")
    return p


def test_cli_citeseer(xgm, tmp_path):
    d = os.path.dirname(CITESEER)
    for graph in (d, _edgelist(xgm, tmp_path)):
        _say("cli", graph)
        assert _counts(xgm, "-g", graph, "--patternID", "20") == {"P4": 255}
        assert _counts(xgm, "-g", graph, "--patternID", "21", "--algorithm", "glumin_g2miner") == {"P5": 46}
        assert _counts(xgm, "-g", graph, "--pattern-name", "triangle", "--repeat", "2") == {"triangle": 1166}
        assert _counts(xgm, "-g", graph, "--pattern-name", "diamond") == {"diamond": 3730}
        assert _counts(xgm, "-g", graph, "--pattern-name", "rectangle", "--algorithm", "graphminer") == \
            {"rectangle": 6059}
        assert _counts(xgm, "-g", graph, "--pattern-size", "4", "--pattern-adj-mat", K4).popitem()[1] == 255
        assert _counts(xgm, "-g", graph, "--pattern-size", "4", "--pattern-adj-mat", DIAMOND).popitem()[1] == 3730
        assert _counts(xgm, "-g", graph, "--pattern-size", "4", "--pattern-adj-mat", CYCLE4).popitem()[1] == 6059
    # GLUMIN P1-P3 in unique-embedding semantics (src/motif/README.md:53-56)
    assert _counts(xgm, "-g", d, "--patternID", "17") == {"P1": 222630}
    assert _counts(xgm, "-g", d, "--patternID", "18") == {"P2": 2200}
    assert _counts(xgm, "-g", d, "--patternID", "19") == {"P3": 22900}
    # the reference's X-GMiner path: its bucket-bitmap P2 (main.cpp:48-52)
    assert _counts(xgm, "-g", d, "--patternID", "18", "--run-xgminer", "1") == {"P2": 2200}
    for st in ("lut", "list"):
        assert _counts(xgm, "-g", d, "--patternID", "17", "--strategy", st) == {"P1": 222630}, st
    m = _counts(xgm, "-g", d, "--pattern-name", "motif4")
    assert m == {"3-star": 222630, "4-path": 111153, "tailed-triangle": 22900, "4-cycle": 3094,
                 "diamond": 2200, "4-clique": 255}
