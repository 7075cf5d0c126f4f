"""Train the clique cost model on the GPU and save it.

    python x-gminer_amd/train_scheduler.py [--out PATH] [--quick]

Training graphs: R-MAT scales 20-22 and Chung-Lu graphs (seeds 10-40), each
degree-oriented; the held-out evaluation graph (R-MAT-23 seed 4) is never
seen here.  Prints one JSON line per stage."""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _pkg():
    if "xgminer_amd" in sys.modules:
        return sys.modules["xgminer_amd"]
    spec = importlib.util.spec_from_file_location("xgminer_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["xgminer_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


TRAIN = [("rmat", 20, 10), ("rmat", 21, 20), ("rmat", 22, 30), ("cl", 1000000, 40)]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--quick", action="store_true", help="small graphs (scales 16-18) for tests")
    ap.add_argument("--epochs", type=int, default=60)
    args = ap.parse_args()
    xgm = _pkg()
    from xgminer_amd import scheduler
    graphs = [("rmat", 16, 10), ("rmat", 17, 20), ("cl", 100000, 40)] if args.quick else TRAIN
    ctx = xgm.Context(0)
    Xs, Ts, Ms = [], [], []
    for kind, size, seed in graphs:
        t0 = time.time()
        if kind == "rmat":
            g = xgm.gen_rmat(size, 16, 0.57, 0.19, 0.19, seed)
        else:
            g = xgm.gen_chung_lu(size, 20.0, 2.1, 5000.0 if size < 1e6 else 50000.0, seed)
        d = g.orient()
        dg = ctx.upload(d)
        codes = range(5) if (kind != "rmat" or size <= 21) else (0, 1, 2, 4)  # edge lists are slow at scale 22
        X, T, M, cnt = scheduler.collect(dg, xgm.CLIQUE5, codes=codes, max_rows=300000, seed=seed)
        dg.close()
        Xs.append(X); Ts.append(T); Ms.append(M)
        print(json.dumps(dict(stage="collect", graph=f"{kind}{size}_s{seed}", tasks=int(len(X)), k5=cnt,
                              seconds=round(time.time() - t0, 1))), flush=True)
    X, T, M = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)
    t0 = time.time()
    model, loss = scheduler.train(X, T, M, epochs=args.epochs)
    pred = model.predict(X)
    choice = np.argmin(np.where(M > 0, pred, np.inf), axis=1)
    best = scheduler.oracle_best(T, M)
    cost = lambda ch: float(np.sum(2.0 ** T[np.arange(len(T)), ch] - 1.0) * 1e-6)  # ms of kernel time
    print(json.dumps(dict(stage="train", rows=int(len(X)), final_loss=loss, seconds=round(time.time() - t0, 1),
                          train_cost_ms=dict(model=cost(choice), best=cost(best),
                                             **{f"code{c}": cost(np.where(M[:, c] > 0, c, 2)) for c in range(5)}))),
          flush=True)
    out = args.out or scheduler.DEFAULT_MODEL
    scheduler.save(model, out)
    print(json.dumps(dict(stage="saved", path=out)), flush=True)


if __name__ == "__main__":
    main()
