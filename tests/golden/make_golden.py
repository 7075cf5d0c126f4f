"""Writes tests/golden/counts.json: exact counts of seeded graphs computed by
the reference's own GraphMiner OpenMP binaries (oracle/_ref, built from the
reference sources by oracle/Makefile), keyed by the graph's content hash.

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden.py [--full]
--full adds the full-size benchmark graphs (minutes to hours of CPU time).
The graphs themselves are regenerated by the tests from (generator, params);
only their hash, size and counts are stored here."""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from conftest import load_xgm  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "counts.json")

# name -> (generator, kwargs, patterns)
SMALL = {
    "rmat10_s7": ("rmat", dict(scale=10, edge_factor=16, seed=7), "all"),
    "rmat12_s3": ("rmat", dict(scale=12, edge_factor=16, seed=3), "all"),
    "er2000_d16_s5": ("er", dict(n=2000, avg_degree=16.0, seed=5), "all"),
    "chunglu20k_s9": ("cl", dict(n=20000, avg_degree=10.0, beta=2.1, w_max=500.0, seed=9), "all"),
    "rmat16_s1": ("rmat", dict(scale=16, edge_factor=16, seed=1), "all"),
}
FULL = {
    "rmat22_s2": ("rmat", dict(scale=22, edge_factor=16, seed=2), ["triangle", "clique4", "clique5"]),
    "chunglu4m_s3": ("cl", dict(n=4000000, avg_degree=20.0, beta=2.1, w_max=50000.0, seed=3),
                     ["diamond", "rectangle"]),
    "rmat23_s4": ("rmat", dict(scale=23, edge_factor=16, seed=4), ["clique5"]),
}
ALL = ["triangle", "clique4", "clique5", "diamond", "rectangle"]


def make_graph(xgm, kind, kw):
    if kind == "rmat":
        return xgm.gen_rmat(kw["scale"], kw["edge_factor"], 0.57, 0.19, 0.19, kw["seed"])
    if kind == "er":
        return xgm.gen_erdos_renyi(kw["n"], kw["avg_degree"], kw["seed"])
    if kind == "cl":
        return xgm.gen_chung_lu(kw["n"], kw["avg_degree"], kw["beta"], kw["w_max"], kw["seed"])
    raise ValueError(kind)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    xgm = load_xgm()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    todo = dict(SMALL)
    if args.full:
        todo.update(FULL)
    if args.only:
        todo = {k: v for k, v in todo.items() if k in args.only.split(",")}
    for name, (kind, kw, pats) in todo.items():
        g = make_graph(xgm, kind, kw)
        d = g.orient()
        ent = data.get(name, {})
        ent.update(dict(generator=kind, params=kw, hash=f"{g.hash():016x}", dag_hash=f"{d.hash():016x}",
                        n=g.n, undirected_edges=g.nnz // 2, dag_max_out_degree=int(d.degree().max())))
        ent.setdefault("counts", {})
        ent.setdefault("ref_seconds", {})
        for p in (ALL if pats == "all" else pats):
            if p in ent["counts"]:
                continue
            t0 = time.time()
            c, t = oracle.ref_count_graph(p, g, timeout=6 * 3600)
            ent["counts"][p] = c
            ent["ref_seconds"][p] = t
            print(f"{name} {p} = {c} (ref loop {t:.2f}s, wall {time.time()-t0:.1f}s)", flush=True)
            data[name] = ent
            json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)
        data[name] = ent
        json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
