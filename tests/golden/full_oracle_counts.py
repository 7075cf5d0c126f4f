"""Full-size golden counts with the oracle (oracle/oracle.cpp, itself pinned
against the reference's GraphMiner binaries), computed in vertex-range
chunks so that long counts report progress.  Writes a JSON fragment that is
merged into tests/golden/counts.json with its source.

    python tests/golden/full_oracle_counts.py chunglu4m_s3 diamond,rectangle OUT.json [threads] [chunks] [method]

method "restatement" (default) runs the GraphMiner restatement; "bitset"
counts 4/5-cliques with the oracle's bitset sub-DAG counter
(orc_clique_bitset, pinned against the restatement in tests/test_oracle.py).
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from conftest import load_xgm  # noqa: E402
import oracle  # noqa: E402
from golden.make_golden import FULL, SMALL, make_graph  # noqa: E402


def main():
    name, pats, out = sys.argv[1], sys.argv[2].split(","), sys.argv[3]
    threads = int(sys.argv[4]) if len(sys.argv) > 4 else 16
    chunks = int(sys.argv[5]) if len(sys.argv) > 5 else 64
    method = sys.argv[6] if len(sys.argv) > 6 else "restatement"
    xgm = load_xgm()
    kind, kw, _ = {**SMALL, **FULL}[name]
    t0 = time.time()
    g = make_graph(xgm, kind, kw)
    res = dict(name=name, hash=f"{g.hash():016x}", counts={}, seconds={}, threads=threads,
               source=("oracle bitset clique counter (orc_clique_bitset)" if method == "bitset"
                       else "oracle restatement (oracle/oracle.cpp)") + ", vertex-range chunks")
    print(json.dumps(dict(stage="graph", n=g.n, hash=res["hash"], s=round(time.time() - t0, 1))), flush=True)
    for p in pats:
        d = g.orient() if p in ("triangle", "clique4", "clique5") else g
        total, B = 0, 0
        t1 = time.time()
        step = (d.n + chunks - 1) // chunks
        for c in range(chunks):
            lo, hi = c * step, min(d.n, (c + 1) * step)
            if method == "bitset":
                cnt, b = oracle.clique_bitset(d.rowptr, d.colidx, int(p[-1]), lo, hi, threads), 0
            else:
                cnt, b = oracle.count(p, d.rowptr, d.colidx, lo, hi, threads)
            total += cnt
            B += b
            print(json.dumps(dict(stage="chunk", pattern=p, chunk=c, lo=lo, hi=hi, partial=total,
                                  s=round(time.time() - t1, 1))), flush=True)
        res["counts"][p] = total
        res["seconds"][p] = round(time.time() - t1, 1)
        res.setdefault("alg_bytes", {})[p] = B
        json.dump(res, open(out, "w"), indent=1)
        print(json.dumps(dict(stage="done", pattern=p, count=total, seconds=res["seconds"][p])), flush=True)


if __name__ == "__main__":
    main()
