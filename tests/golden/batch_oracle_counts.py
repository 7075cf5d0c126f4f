"""Golden counts of the 256-instance batch (x-gminer_amd/workloads.py):
triangles and 4-cliques on each instance's degree-oriented DAG, edge-induced
diamonds and 4-cycles on the undirected instance, computed one instance at a
time with the oracle (oracle/oracle.cpp, pinned against the reference's
GraphMiner binaries).  4-cycles use the rectangle.h restatement
("rectangle_method": "restatement") or, where that takes tens of minutes
per instance, the oracle's wedge-aggregation counter ("wedge"), which
tests/test_oracle.py pins against the restatement.  Writes
tests/golden/batch_counts.json: per instance the seed, kind, parameters,
CSR hash and the four counts.

    python tests/golden/batch_oracle_counts.py [threads] [first_n_per_half]
    python tests/golden/batch_oracle_counts.py --recheck [threads]
        (recount every "restatement" 4-cycle with the wedge counter)
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

OUT = os.path.join(ROOT, "tests", "golden", "batch_counts.json")
PATTERNS = ("triangle", "clique4", "diamond", "rectangle")


def recheck(threads):
    xgm = load_xgm()
    from xgminer_amd import workloads
    data = json.load(open(OUT))
    for r in data["instances"]:
        if r.get("rectangle_method", "restatement") != "restatement":
            continue
        g = workloads.make_instance(xgm, (r["seed"], r["kind"], r["params"]))
        c = oracle.cycle4_wedge(g.rowptr, g.colidx, threads)
        assert c == r["counts"]["rectangle"], r["seed"]
        r["rectangle_wedge_agrees"] = True
        print(json.dumps(dict(seed=r["seed"], rectangle=c)), flush=True)
    json.dump(data, open(OUT, "w"), indent=0)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--recheck":
        return recheck(int(sys.argv[2]) if len(sys.argv) > 2 else 8)
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    count = int(sys.argv[2]) if len(sys.argv) > 2 else None
    xgm = load_xgm()
    from xgminer_amd import workloads
    specs = workloads.batch_specs(count)
    done = {}
    if os.path.exists(OUT):
        for r in json.load(open(OUT))["instances"]:
            done[r["seed"]] = r
    rows = []
    t0 = time.time()
    for spec in specs:
        seed, kind, p = spec
        if seed in done:
            rows.append(done[seed])
            continue
        g = workloads.make_instance(xgm, spec)
        d = g.orient()
        row = dict(seed=seed, kind=kind, params=p, n=g.n, nnz=g.nnz, hash=f"{g.hash():016x}", counts={},
                   rectangle_method="wedge")
        for pat in PATTERNS[:3]:
            h = d if pat in ("triangle", "clique4") else g
            row["counts"][pat] = oracle.count(pat, h.rowptr, h.colidx, 0, -1, threads)[0]
        row["counts"]["rectangle"] = oracle.cycle4_wedge(g.rowptr, g.colidx, threads)
        rows.append(row)
        print(json.dumps(dict(seed=seed, n=g.n, nnz=g.nnz, counts=row["counts"], s=round(time.time() - t0, 1))),
              flush=True)
        json.dump(dict(source="oracle (oracle/oracle.cpp), one instance at a time",
                       instances=sorted(rows + [r for s, r in done.items() if s not in {x["seed"] for x in rows}],
                                        key=lambda r: r["seed"])), open(OUT, "w"), indent=0)


if __name__ == "__main__":
    main()
