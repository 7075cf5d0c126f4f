"""Benchmark of every configuration on one MI355X.

Default run (no --workload): the whole suite in one JSON line.
* value: |E|/s of triangle counting on R-MAT scale 22 (edge factor 16,
  (a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2, degree-oriented int32 CSR); a
  step is one full count over the resident graph.
* per_pattern: triangles on R-MAT-16; 4- and 5-cliques on R-MAT-22 (the
  library default, the reference-style id-LUT threshold, the learned
  schedule); edge-induced diamonds and 4-cycles on Chung-Lu 4M; the
  256-instance batch (one hipGraph replay of {triangle, 4-clique, diamond,
  4-cycle} on every instance); 5-cliques on R-MAT-23 (learned against the
  model-free rank-LUT schedule).  Every count is checked against its golden
  count.
* every row: HIP-event kernel times, the dominant kernel's roofline under a
  byte model of its own algorithm (each list it reads, once), its HBM
  traffic from two rocprofv3 --pmc runs, and a CPU baseline (the reference
  GraphMiner binaries, or the oracle's restatement on a task sample where
  they take hours) on the host's CPU share with OMP_PROC_BIND=spread.

--workload / --pattern / --strategy / --sched / --prep select the single
modes (one workload, one pattern).  Progress goes to stderr."""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_CLOCK_GHZ = 2.4  # MI355X peak engine clock
COPY_BYTES = 1 << 30  # copy kernel size for bandwidth and counter calibration
PATTERN_IDS = {"triangle": 1, "clique4": 2, "clique5": 3, "diamond": 4, "rectangle": 5}
# workload name -> (generator, params, golden key, graph the patterns run on)
WORKLOADS = {
    "rmat22": ("rmat", dict(scale=22, seed=2), "rmat22_s2", "dag"),
    "rmat16": ("rmat", dict(scale=16, seed=1), "rmat16_s1", "dag"),
    "rmat23": ("rmat", dict(scale=23, seed=4), "rmat23_s4", "dag"),
    "chunglu": ("cl", dict(n=4000000, avg_degree=20.0, beta=2.1, w_max=50000.0, seed=3), "chunglu4m_s3",
                "undirected"),
    "batch256": ("batch", {}, None, "both"),
}
BATCH_PATTERNS = ("triangle", "clique4", "diamond", "rectangle")
DIAMOND_TIERS = [("diamond_wave_64", 1, 64), ("diamond_block_1024", 65, 1024), ("diamond_block_8192", 1025, 8192),
                 ("diamond_hbm_set", 8193, 1 << 40)]
# tier bounds of the triangle kernels (out-degree) under the library's default
# split (kernels.h: kTcWaveCap, kTcWaveLargeDefault = 64, so the 256-slot wave
# tier is empty and the block kernel takes 65..2048)
TC_TIERS = [("tc_wave_64", 1, 64), ("tc_block_2048", 65, 2048), ("tc_block_big", 2049, 1 << 40)]


_T0 = time.time()


def log(*a):
    print(f"[bench {time.time() - _T0:6.1f}s]", *a, file=sys.stderr, flush=True)


def tc_alg_bytes_per_tier(d):
    """Algorithmic bytes of triangle counting, 4*(d+(u) + d+(v)) + 40 per DAG
    edge u->v (8 B src/dst + 32 B rowptr), attributed to the kernel that
    processes the edge: the library charges an edge to the endpoint with the
    longer out-list (ties to v) and tiers items by that out-degree."""
    dout = np.diff(d.rowptr).astype(np.int64)
    src = np.repeat(np.arange(d.n, dtype=np.int64), dout)
    du = dout[src]
    dv = dout[d.colidx]
    del src
    t_deg = np.where(du <= dv, dv, du)
    per_e = 4 * (du + dv) + 40
    out = {}
    for name, lo, hi in TC_TIERS:
        m = (t_deg >= lo) & (t_deg <= hi)
        out[name] = int(per_e[m].sum())
    return out, int(per_e.sum())


def diamond_alg_bytes_per_tier(g):
    """Algorithmic bytes of edge-induced diamond counting on an undirected
    CSR: 4*(d(v0) + d(v1)) + 40 per undirected edge, attributed to the kernel
    tier of its owner (the (degree, id)-larger endpoint)."""
    deg = np.diff(g.rowptr).astype(np.int64)
    src = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    dst = g.colidx.astype(np.int64)
    keep = src < dst
    src, dst = src[keep], dst[keep]
    ds, dd = deg[src], deg[dst]
    owner_deg = np.maximum(ds, dd)
    per_e = 4 * (ds + dd) + 40
    out = {name: int(per_e[(owner_deg >= lo) & (owner_deg <= hi)].sum()) for name, lo, hi in DIAMOND_TIERS}
    return out, int(per_e.sum())


def union(xgm, graphs):
    """Disjoint union of host graphs (the layout xgm_graph_upload_batch builds)."""
    offs = np.cumsum([0] + [g.n for g in graphs])
    eoffs = np.cumsum([0] + [g.nnz for g in graphs])
    rp = np.concatenate([g.rowptr[:-1] + e for g, e in zip(graphs, eoffs)] + [np.array([eoffs[-1]])])
    ci = np.concatenate([g.colidx + np.int32(o) for g, o in zip(graphs, offs)])
    return xgm.Graph.from_arrays(rp.astype(np.int64), ci.astype(np.int32), graphs[0].flags)


def cpu_baseline_batch(specs, inst, threads, budget_s=12.0):
    """Reference GraphMiner OpenMP binaries on a bounded sample of the batch:
    instances at evenly spaced size quantiles (8 of them, smallest first),
    each counted for all four patterns; an instance whose predicted time
    (its edges over the rate measured so far) exceeds what is left of
    budget_s counting-loop seconds is skipped.  |E|/s = sum over the sampled
    instances of 4 m / counting-loop seconds (the batch's GPU unit)."""
    import oracle
    if not all(oracle.ref_available(p) for p in BATCH_PATTERNS):
        return None
    by_m = sorted(range(len(inst)), key=lambda i: inst[i].nnz)
    k = 8
    picks = [by_m[(2 * j + 1) * len(by_m) // (2 * k)] for j in range(k)]
    tot_t, tot_e, used, skipped = 0.0, 0, [], []
    for i in picks:
        g = inst[i]
        m = g.nnz // 2
        if tot_t > 0 and (4 * m) / (tot_e / tot_t) > budget_s - tot_t:
            skipped.append(specs[i][0])
            continue
        for p in BATCH_PATTERNS:
            _, t = oracle.ref_count_graph(p, g, threads=threads, timeout=600)
            tot_t += t
            tot_e += m
        used.append(specs[i][0])
        log(f"cpu baseline: batch seed {specs[i][0]} (m={m}) done, {tot_t:.1f}s")
    return dict(value=tot_e / tot_t, unit="|E|/s", cores=threads, kind="reference",
                sample=f"{len(used)} batch instances at size quantiles (seeds {used}; skipped over budget: "
                       f"{skipped}), all four patterns each, GraphMiner tc/clique/sgl OpenMP binaries "
                       "(counting loops only)", seconds=round(tot_t, 3))


def bench_batch(args, xgm):
    from xgminer_amd import workloads
    t0 = time.time()
    specs = workloads.batch_specs()
    inst = workloads.batch_graphs(xgm, specs)
    dags = [g.orient() for g in inst]
    msum = sum(g.nnz // 2 for g in inst)
    log(f"batch: {len(inst)} instances, sum m={msum} ({time.time()-t0:.1f}s)")
    gpath = os.path.join(ROOT, "tests", "golden", "batch_counts.json")
    golden = None
    if os.path.exists(gpath):
        rows = json.load(open(gpath))["instances"]
        if len(rows) == len(specs) and all(r["seed"] == s[0] for r, s in zip(rows, specs)):
            if all(f"{g.hash():016x}" == r["hash"] for g, r in zip(inst, rows)):
                golden = np.array([[r["counts"][p] for r in rows] for p in BATCH_PATTERNS], np.uint64)
    ctx = xgm.Context(0)
    t0 = time.time()
    dag = ctx.upload_batch(dags)
    und = ctx.upload_batch(inst)
    log(f"upload {time.time()-t0:.2f}s")
    pids = [PATTERN_IDS[p] for p in BATCH_PATTERNS]
    on = {"triangle": dag, "clique4": dag, "diamond": und, "rectangle": und}

    t0 = time.time()
    plain = ctx.plan(pids, dag=dag, und=und)
    plain_counts, _ = plain.run()
    log(f"heuristic plan ready ({time.time()-t0:.1f}s)")
    t0 = time.time()
    tuned = {}
    for p in BATCH_PATTERNS:
        best, ms = ctx.autotune(on[p], PATTERN_IDS[p], reps=2)
        tuned[p] = dict(opts=best, ms=round(ms, 3))
        log(f"autotune {p}: {best} {ms:.2f} ms")
    tune_s = time.time() - t0
    plan = ctx.plan(pids, dag=dag, und=und, opts=[tuned[p]["opts"] for p in BATCH_PATTERNS])

    def timed(pl, steps, warmup):
        counts = None
        for _ in range(warmup):
            counts, _ = pl.run()
        evs = []
        t = time.perf_counter()
        for _ in range(steps):
            c, ms = pl.run()
            evs.append(ms)
            if counts is not None and not np.array_equal(c, counts):
                raise SystemExit("non-deterministic batch counts")
            counts = c
        return counts, (time.perf_counter() - t) * 1e3 / steps, float(np.mean(evs))

    counts, ms_step, ev_ms = timed(plan, args.steps, args.warmup)
    pcounts, pms_step, pev_ms = timed(plain, args.steps, args.warmup)
    log(f"tuned plan {ms_step:.3f} ms/step (graph {ev_ms:.3f}); heuristic plan {pms_step:.3f} ms/step")
    if not np.array_equal(counts, pcounts) or not np.array_equal(counts, plain_counts):
        raise SystemExit("tuned and heuristic plans disagree")
    if golden is not None and not np.array_equal(counts, golden):
        bad = np.argwhere(counts != golden)[:5].tolist()
        raise SystemExit(f"BATCH COUNT MISMATCH vs golden at {bad}")

    # per-pattern eager counts with the tuned options: kernel durations (HIP
    # events on the library stream) for the roofline of the dominant kernel
    kernels = {}
    for p in BATCH_PATTERNS:
        ctx.count_instances(on[p], PATTERN_IDS[p], **tuned[p]["opts"])
        _, ks = ctx.last_profile()
        for k in ks:
            kernels[k.name] = kernels.get(k.name, 0.0) + k.ms
    dom = max(kernels, key=kernels.get)
    roof = None
    alg = {}
    if dom.startswith("tc_"):
        alg, _ = tc_alg_bytes_per_tier(union(xgm, dags))
    elif dom.startswith("diamond_"):
        alg, _ = diamond_alg_bytes_per_tier(union(xgm, inst))
    if dom in alg:
        ach = alg[dom] / (kernels[dom] * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel=dom,
                    alg_bytes_per_launch=alg[dom], avg_launch_ms=round(kernels[dom], 4))
    cpu = None if args.no_cpu else cpu_baseline_batch(specs, inst, args.cpu_threads)
    out = {
        "metric": "|E|/s on batch256 (sum of m over instances x 4 patterns per step)",
        "value": 4 * msum / (ms_step * 1e-3),
        "unit": "|E|/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "256 instances (128 Erdos-Renyi n log-uniform [1e3, 2e5], avg degree U[4, 64]; "
                               "128 R-MAT scale U{10..18}, edge factor 16; seeds 100..355) x {triangle, "
                               "4-clique, diamond, 4-cycle}, one hipGraph replay per step",
                   "instances": len(inst), "sum_undirected_edges": msum, "parallelism": "single GPU"},
        "counts_total": {p: int(counts[i].sum()) for i, p in enumerate(BATCH_PATTERNS)},
        "count_matches_golden": True if golden is not None else None,
        "graph_replay_ms": round(ev_ms, 4),
        "heuristic_plan_ms_per_step": round(pms_step, 4),
        "tuned_vs_heuristic": round(pms_step / ms_step, 4),
        "autotune": tuned,
        "autotune_s": round(tune_s, 1),
        "kernels_ms": {k: round(v, 4) for k, v in kernels.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    for x in (plan, plain):
        x.close()
    ctx.close()


# kernel symbol prefix of each profiled launch name (rocprof sees symbols)
KERNEL_SYMBOL = {"tc_block_2048": "k_tc_block<", "tc_block_big": "k_tc_block<", "tc_wave_256": "k_tc_wave<256",
                 "tc_wave_64": "k_tc_wave<64,"}


def _pmc_per_dispatch(counter, argv, symbol, timeout=240):
    """Run `python3 bench.py argv` under rocprofv3 --pmc counter and return
    the counter's mean value per dispatch of kernels whose name contains
    `symbol` (KiB for FETCH_SIZE / WRITE_SIZE), or None."""
    import csv
    import glob
    import shutil
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    with tempfile.TemporaryDirectory(prefix="xgm_pmc_") as d:
        cmd = [exe, "--pmc", counter, "--kernel-trace", "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__)] + argv
        if not _run_logged(cmd, timeout, f"rocprofv3 --pmc {counter}"):
            return None
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        vals = {}
        for f in files:
            for row in csv.DictReader(open(f)):
                if symbol in row.get("Kernel_Name", "") and row.get("Counter_Name") == counter:
                    key = row.get("Dispatch_Id") or row.get("Correlation_Id") or str(len(vals))
                    vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
        if not vals:
            return None
        return float(np.mean(list(vals.values())))


def measure_traffic(args, dom):
    """HBM bytes per launch of kernel `dom` from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE), each scaled by what the same counter reads for
    the plain copy kernel over a known size (bytes / (counter KiB * 1024))."""
    sym = KERNEL_SYMBOL.get(dom)
    if sym is None:
        return None, None
    run = ["--workload", args.workload, "--pattern", args.pattern, "--strategy", str(args.strategy),
           "--steps", "1", "--warmup", "0", "--no-cpu", "--no-traffic"]
    cal = ["--copy-calibrate"]
    out = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        log(f"traffic: rocprofv3 --pmc {c}")
        k = _pmc_per_dispatch(c, run, sym)
        ref = _pmc_per_dispatch(c, cal, "k_copy_u4")
        if k is None or ref is None or ref <= 0:
            return None, None
        factor = COPY_BYTES / (ref * 1024.0)
        out[c] = dict(kib_per_dispatch=round(k, 1), copy_kib=round(ref, 1), factor=round(factor, 4),
                      bytes=k * 1024.0 * factor)
    return int(out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"]), out


def bench_prep(args, xgm, g, m, desc):
    """Undirected host CSR -> device DAG ready to count, two ways: host
    orientation (xgm_orient_dag) then xgm_graph_upload (host task order,
    reverse lists, H2D), against xgm_graph_upload of the undirected CSR then
    xgm_graph_orient (kernels in HBM, host mirrors copied back).  Wall time
    per leg, median over --steps after --warmup; the two DAGs are checked
    equal (CSR hash and task order)."""
    ctx = xgm.Context(0)
    legs = {"host_orient_upload": [], "upload_undirected": [], "device_orient": [], "host_route_first_tc": [],
            "device_route_first_tc": []}
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        d = g.orient()
        dg = ctx.upload(d)
        t1 = time.perf_counter()
        du = ctx.upload(g)
        t2 = time.perf_counter()
        dd = ctx.orient(du)
        t3 = time.perf_counter()
        # first triangle count on each DAG: its task items are built then
        # (on the host from the host DAG, in HBM for the device-built one)
        ch = ctx.count(dg, 1)
        t4 = time.perf_counter()
        cd = ctx.count(dd, 1)
        t5 = time.perf_counter()
        if ch != cd:
            raise SystemExit(f"triangle counts differ between the routes: {ch} vs {cd}")
        if i == 0:
            if dd.download().hash() != d.hash() or not np.array_equal(dd.task_order(), dg.task_order()):
                raise SystemExit("device orientation differs from the host orientation")
        if i >= args.warmup:
            legs["host_orient_upload"].append((t1 - t0) * 1e3)
            legs["upload_undirected"].append((t2 - t1) * 1e3)
            legs["device_orient"].append((t3 - t2) * 1e3)
            legs["host_route_first_tc"].append((t4 - t3) * 1e3)
            legs["device_route_first_tc"].append((t5 - t4) * 1e3)
        log(f"step {i}: host {1e3 * (t1 - t0):.1f} ms, upload {1e3 * (t2 - t1):.1f} ms, "
            f"device orient {1e3 * (t3 - t2):.1f} ms; first count host route {1e3 * (t4 - t3):.1f} ms, "
            f"device route {1e3 * (t5 - t4):.1f} ms")
        for x in (dd, du, dg):
            x.close()
        del d
    med = {k: round(float(np.median(v)), 3) for k, v in legs.items()}
    dev_total = med["upload_undirected"] + med["device_orient"]
    to_first = {"host_route": round(med["host_orient_upload"] + med["host_route_first_tc"], 3),
                "device_route": round(dev_total + med["device_route_first_tc"], 3)}
    print(json.dumps({
        "metric": f"host CSR to device DAG on {args.workload}, ms (orientation, reverse lists, task order)",
        "value": round(dev_total, 3), "unit": "ms", "higher_is_better": False, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "config": {"workload": desc, "n": g.n, "undirected_edges": m},
        "legs_ms": med, "host_csr_to_first_triangle_count_ms": to_first, "host_path_ms": med["host_orient_upload"],
        "speedup_vs_host_path": round(med["host_orient_upload"] / dev_total, 3),
        "host_threads": int(os.environ.get("XGM_NUM_THREADS", 0)) or (os.cpu_count() or 1),
    }), flush=True)
    ctx.close()


def bench_prep_undirected(args, xgm, g, m, desc):
    """First counts of the undirected patterns (diamonds, 4-cycles, the
    4-motif census) from a fresh upload, with the task lists built on the
    host (XGM_PREP_HOST: the CSR copied back, the host builders) and in HBM
    (XGM_PREP_DEVICE), beside the next count, which finds them cached.  Wall
    time per leg, median over --steps after --warmup; both routes must give
    the same counts."""
    ctx = xgm.Context(0)
    pats = [("diamond", lambda dg: ctx.count(dg, PATTERN_IDS["diamond"])),
            ("rectangle", lambda dg: ctx.count(dg, PATTERN_IDS["rectangle"])),
            ("motif4", lambda dg: tuple(ctx.motifs(dg, 4)[0].tolist()))]
    legs = {}
    for i in range(args.warmup + args.steps):
        for name, f in pats:
            res = {}
            for route, rname in ((xgm.PREP_HOST, "host"), (xgm.PREP_DEVICE, "device")):
                ctx.set_prep(route)
                dg = ctx.upload(g)
                t0 = time.perf_counter()
                res[rname] = f(dg)
                t1 = time.perf_counter()
                f(dg)
                t2 = time.perf_counter()
                dg.close()
                log(f"step {i} {name} {rname} route: first {1e3 * (t1 - t0):.1f} ms, cached {1e3 * (t2 - t1):.1f} ms")
                if i >= args.warmup:
                    legs.setdefault(f"{name}_{rname}_route_first", []).append((t1 - t0) * 1e3)
                    legs.setdefault(f"{name}_cached", []).append((t2 - t1) * 1e3)
            if res["host"] != res["device"]:
                raise SystemExit(f"{name}: the routes differ: {res['host']} vs {res['device']}")
    ctx.set_prep(xgm.PREP_AUTO)
    med = {k: round(float(np.median(v)), 3) for k, v in legs.items()}
    print(json.dumps({
        "metric": f"first count of the undirected patterns on {args.workload}, ms (task lists built in HBM)",
        "value": med["rectangle_device_route_first"], "unit": "ms", "higher_is_better": False, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "config": {"workload": desc, "n": g.n, "undirected_edges": m},
        "legs_ms": med,
        "host_threads": int(os.environ.get("XGM_NUM_THREADS", 0)) or (os.cpu_count() or 1),
    }), flush=True)
    ctx.close()


def bench_motif(args, xgm, g, m, desc, gold_counts):
    """Vertex-induced 3- or 4-motif census of the undirected workload graph
    (xgm_motifs: the reference's formula path).  Wall time per census over
    --steps after --warmup (the oriented copy and task lists are built in the
    warm-up).  Checks: the census' triangle (k = 3) or 4-clique (k = 4) entry
    against the golden count of the graph.  CPU baseline: the reference's
    motif_omp_formula binary on R-MAT-16 (seed 1), a bounded sample."""
    k = int(args.pattern[-1])
    ctx = xgm.Context(0)
    dg = ctx.upload(g)
    census = None
    for i in range(max(1, args.warmup)):
        census = ctx.motifs(dg, k)[0].tolist()
        log(f"warmup {i}: {census}")
    t0 = time.perf_counter()
    for i in range(args.steps):
        c = ctx.motifs(dg, k)[0].tolist()
        if c != census:
            raise SystemExit(f"non-deterministic census: {c} vs {census}")
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    check_key, check_idx = ("triangle", 1) if k == 3 else ("clique4", 5)
    want = gold_counts.get(check_key)
    if want is not None and census[check_idx] != want:
        raise SystemExit(f"COUNT MISMATCH vs golden {check_key}: {census[check_idx]} != {want}")
    cpu = None
    if not args.no_cpu:
        import oracle
        if oracle.ref_available("triangle") and os.access(os.path.join(oracle.REF_DIR, "motif_omp_formula"), os.X_OK):
            s = xgm.gen_rmat(16, 16, 0.57, 0.19, 0.19, 1)
            log(f"cpu baseline: reference motif_omp_formula on R-MAT-16, {args.cpu_threads} threads")
            ref, sec = oracle.ref_motif_timed(s, k, threads=args.cpu_threads, timeout=300)
            sd = ctx.upload(s)
            ok = ctx.motifs(sd, k)[0].tolist() == ref
            sd.close()
            cpu = dict(value=(s.nnz // 2) / sec, unit="|E|/s", cores=args.cpu_threads, kind="reference",
                       sample=f"R-MAT-16 seed 1 ({s.nnz // 2} edges), full {k}-motif census (counting loop only)",
                       seconds=sec, census_matches_gpu=ok)
    print(json.dumps({
        "metric": f"|E|/s on {args.workload} ({k}-motif census)", "value": m / (ms * 1e-3), "unit": "|E|/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": f"vertex-induced {k}-motif census on {desc}", "n": g.n, "undirected_edges": m},
        "census": dict(zip(oracle_motif_names(k), census)),
        "golden_check": {check_key: want, "matches": (want is None) or census[check_idx] == want},
        "roofline": None, "cpu_baseline": cpu,
    }), flush=True)
    ctx.close()


def oracle_motif_names(k):
    return ["wedge", "triangle"] if k == 3 else ["3-star", "4-path", "tailed-triangle", "4-cycle", "diamond",
                                                  "4-clique"]


def bench_sched(args, xgm, d, m, desc, golden):
    """Learned vs heuristic clique schedules on one graph: the learned
    schedule is timed over the K steps; the heuristic (and the all-rank-LUT
    schedule) over --baseline-steps steps, every count checked equal."""
    from xgminer_amd import scheduler
    pid = PATTERN_IDS[args.pattern]
    ctx = xgm.Context(0)
    dg = ctx.upload(d)
    t0 = time.time()
    model = scheduler.load()
    learned = scheduler.schedule(dg, model, pid)
    prep_s = time.time() - t0
    plans = {"learned": learned, "heuristic": dg.sched_heuristic(pid), "rank_all": dg.sched_heuristic(pid, 6)}
    res, count = {}, None
    for name, codes in plans.items():
        s = xgm.Schedule(dg, pid, codes)
        steps = args.steps if name == "learned" else args.baseline_steps
        for i in range(args.warmup if name == "learned" else 0):
            s.count()
        times, t = [], time.perf_counter()
        for i in range(steps):
            c = s.count()
            times.append(ctx.last_profile()[0])
            log(f"{name} step {i}: count={c} kernel ms={times[-1]:.1f}")
            if count is not None and c != count:
                raise SystemExit(f"schedules disagree: {name} {c} vs {count}")
            count = c
        wall_ms = (time.perf_counter() - t) * 1e3 / steps
        res[name] = dict(ms_per_step=round(wall_ms, 3), kernel_ms=round(float(np.median(times)), 3),
                         code_hist=np.bincount(codes[codes >= 0], minlength=5).tolist())
        s.close()
    if golden is not None and count != golden:
        raise SystemExit(f"COUNT MISMATCH vs golden: {count} != {golden}")
    ms = res["learned"]["ms_per_step"]
    out = {
        "metric": f"|E|/s on {args.workload} ({args.pattern}, learned schedule)",
        "value": m / (ms * 1e-3),
        "unit": "|E|/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.pattern} count on {desc}, degree-oriented DAG; learned vs heuristic schedule",
                   "n": d.n, "undirected_edges": m, "parallelism": "single GPU"},
        "count": count,
        "count_matches_golden": (golden == count) if golden is not None else None,
        "speedup_vs_heuristic": round(res["heuristic"]["ms_per_step"] / ms, 3),
        "speedup_vs_rank_all": round(res["rank_all"]["ms_per_step"] / ms, 3),
        "schedules": res,
        "schedule_prep_s": round(prep_s, 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    dg.close()
    ctx.close()


# ------------------------------------------------------------ CPU baselines

def cpu_info():
    """lscpu model and core / thread counts of this host, and the CPUs this
    process may run on."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    if isinstance(info.get("sockets"), int) and isinstance(info.get("cores_per_socket"), int):
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count() or 1
    return info


def cpu_threads(info, requested=0):
    """OpenMP threads of the CPU legs: one per physical core of this
    process's CPU share.  A GPU box here gives one GPU a share of 16 host
    threads (OMP_NUM_THREADS is set to it), whatever lscpu counts."""
    if requested > 0:
        return requested
    share = int(os.environ.get("OMP_NUM_THREADS") or 16)
    tpc = info.get("threads_per_core") if isinstance(info.get("threads_per_core"), int) else 1
    phys = max(1, info.get("affinity_cpus", 1) // max(1, tpc))
    return max(1, min(share, phys, 16))


def set_omp_binding():
    """OMP_PROC_BIND=spread, OMP_PLACES=cores for every CPU leg (the
    in-process oracle reads them when its OpenMP runtime starts; the
    reference binaries inherit them)."""
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")


def ref_timed(pattern, g, threads, single_over_s=8.0, timeout=600):
    """The reference GraphMiner binary on the full graph: median of 3 runs
    of its counting loop (1 run when the first takes longer than
    single_over_s).  Returns (count, median seconds, [seconds])."""
    import oracle
    runs, count = [], None
    with tempfile.TemporaryDirectory(prefix="xgm_ref_") as d:
        prefix = os.path.join(d, "graph")
        g.save(prefix)
        for i in range(3):
            c, t = oracle.ref_count(pattern, prefix, threads=threads, timeout=timeout)
            if count is not None and c != count:
                raise SystemExit(f"reference binary counts differ between runs: {c} vs {count}")
            count = c
            runs.append(t)
            if i == 0 and t > single_over_s:
                break
    return count, float(np.median(runs)), [round(x, 4) for x in runs]


def port_sampled(pattern, gr, order, threads, target_s=6.0, max_pilot_s=1.0):
    """The oracle restatement (GraphMiner's loops) on a stratified sample of
    the vertex tasks: every S-th task of `order` (tasks sorted by decreasing
    work), S a power of two found by a pilot that doubles the sample until it
    runs max_pilot_s, then chosen so the sample runs about target_s.  The
    full-graph time is estimated as sample seconds x len(order) / sample size.
    Returns dict(seconds_est, sample_seconds, fraction, stride, tasks)."""
    import oracle
    order = np.asarray(order, np.int32)
    n = len(order)

    def run(stride):
        tasks = order[stride // 2::stride] if stride > 1 else order
        t0 = time.perf_counter()
        oracle.count_tasks(pattern, gr.rowptr, gr.colidx, tasks, threads)
        dt = time.perf_counter() - t0
        log(f"  {pattern} sample: every {stride}-th of {n} tasks, {dt:.2f}s")
        return dt, len(tasks)

    stride = 1
    while stride * 2 <= n // 64:
        stride *= 2
    t, k = run(stride)
    while t < max_pilot_s and stride > 1:
        stride //= 2
        t, k = run(stride)
    if stride > 1 and t < target_s:
        est_full = t * n / max(1, k)
        want = max(1, int(2 ** math.floor(math.log2(max(1.0, est_full / target_s)))))
        if want < stride:
            stride = want
            t, k = run(stride)
    return dict(seconds_est=t * n / max(1, k), sample_seconds=round(t, 3), fraction=round(k / max(1, n), 6),
                stride=stride, tasks=k)


# ------------------------------------------------------------ byte models
#
# "Bytes the kernel's algorithm must read from HBM" (each list it reads,
# once), per kernel launch name; the roofline fraction of a row is that
# figure over the kernel's HIP-event duration against 8 TB/s.  The
# per-kernel split mirrors the library's task assignment
# (x-gminer_amd/csrc/xgm_api.hip).

def tc_kernel_bytes(d):
    """Triangle kernels: every DAG edge is charged to the endpoint t with the
    longer out-list (ties to the head); t's items stage N+(t) once (4 d+(t)
    + 32 B of row pointers) and read, per charged edge, the partner id (4 B),
    its row pointers (16 B) and its out-list (4 d+(p)).  Kernel by d+(t)."""
    dout = np.diff(d.rowptr).astype(np.int64)
    src = np.repeat(np.arange(d.n, dtype=np.int64), dout)
    du = dout[src]
    dv = dout[d.colidx]
    to_u = du > dv
    t = np.where(to_u, src, d.colidx.astype(np.int64))
    dp = np.where(to_u, dv, du)
    del src
    dt = dout[t]
    per_edge = 20 + 4 * dp
    staged = np.zeros(d.n, bool)
    staged[t] = True
    out = {}
    for name, lo, hi in TC_TIERS:
        me = (dt >= lo) & (dt <= hi)
        mv = staged & (dout >= lo) & (dout <= hi)
        out[name] = int(per_edge[me].sum()) + int((32 + 4 * dout[mv]).sum())
    return out


def clique_kernel_bytes(d, codes, k):
    """Clique kernels: a vertex task v0 reads N+(v0) (16 + 4 d+(v0)) and, for
    every u in N+(v0), its row pointers and out-list (16 + 4 d+(u)); the rank
    LUT also reads each member's in-list bounds (16 B) for its degree.  An
    edge item (v0, v1) reads both out-lists (32 + 4 (d+(v0) + d+(v1))).
    Kernel by schedule code and d+ bin (xgm_api.hip build_clique_sched)."""
    dout = np.diff(d.rowptr).astype(np.int64)
    src = np.repeat(np.arange(d.n, dtype=np.int64), dout)
    s2 = np.bincount(src, weights=dout[d.colidx], minlength=d.n).astype(np.int64)
    B = 16 + 4 * dout + 16 * dout + 4 * s2
    codes = np.asarray(codes)
    task = (dout >= k - 1) & (codes >= 0)
    c = np.where((codes < 0) | (codes > 4), 2, codes)
    hub = task & (dout > 1024) & ((c == 1) | (c == 4)) & (dout <= 4096)
    c = np.where((dout > 1024) & ~hub & np.isin(c, (0, 1, 4)), 2, c)
    binidx = np.searchsorted(np.array([64, 256, 512, 1024]), dout)  # 0..3 for d+ <= 1024
    pre = f"k{k}_"
    out = {}

    def add(name, v):
        out[name] = out.get(name, 0) + int(v)

    caps = (64, 256, 512, 1024)
    for b in range(4):
        m0 = task & (c == 0) & (binidx == b)
        add(pre + f"lut_{caps[b]}", B[m0].sum())
        mr = task & (((c == 1) & (binidx == b)) | ((c == 4) & (np.minimum(binidx + 1, 3) == b))) & ~hub
        add(pre + f"rlut_{caps[b]}", (B[mr] + 16 * dout[mr]).sum())
    tri = hub & (dout <= 1344)
    add(pre + "rlut_tri", (B[tri] + 16 * dout[tri]).sum())
    add(pre + "rlut_hbm", (B[hub & ~tri] + 16 * dout[hub & ~tri]).sum())
    edge = task & ((c == 2) | (c == 3))
    if edge.any():
        dst = d.colidx.astype(np.int64)
        me = edge[src] & (dout[dst] >= k - 2)
        lo = np.minimum(dout[src], dout[dst])[me]
        eb = 32 + 4 * (dout[src][me] + dout[dst][me])
        kind = np.where(c[src][me] == 2, "edge_lut", "edge_list")
        for kk in ("edge_lut", "edge_list"):
            for cap in (128, 1024):
                sel = (kind == kk) & ((lo <= 128) == (cap == 128))
                add(pre + f"{kk}_{cap}", eb[sel].sum())
    return {kname: v for kname, v in out.items() if v}


def diamond_kernel_bytes(g):
    """Diamond kernels: every undirected edge is owned by its (degree,
    id)-larger endpoint t (owners of degree <= 2 hold no diamond and are
    skipped); t's items stage N(t) once (16 + 4 d(t)) and read, per owned
    edge, the partner id, row pointers and list (20 + 4 d(p)).  Kernel by
    d(t)."""
    deg = np.diff(g.rowptr).astype(np.int64)
    src = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    dst = g.colidx.astype(np.int64)
    ds, dd = deg[src], deg[dst]
    own = (ds > dd) | ((ds == dd) & (src > dst))  # src owns the edge
    t, p = src[own], dst[own]
    del src, dst
    dt = deg[t]
    keep = dt >= 3
    per_edge = 20 + 4 * deg[p]
    staged = np.zeros(g.n, bool)
    staged[t[keep]] = True
    out = {}
    for name, lo, hi in DIAMOND_TIERS:
        me = keep & (dt >= lo) & (dt <= hi)
        mv = staged & (deg >= lo) & (deg <= hi)
        out[name] = int(per_edge[me].sum()) + int((16 + 4 * deg[mv]).sum())
    return out


def cycle4_bytes(g):
    """4-cycle wedge aggregation (all kernels of the count): on the (degree,
    id)-relabelled graph every task x reads its lower part of N(x) (16 + 4
    nlow(x)) and, per lower neighbour v, v's row pointers and the part of
    N(v) below x (16 + 4 pos_v(x)).  Summed over x, the part of N(v) read is
    C(d(v), 2) - C(L(v), 2) ids, L(v) = neighbours of v below v."""
    deg = np.diff(g.rowptr).astype(np.int64)
    src = np.repeat(np.arange(g.n, dtype=np.int64), deg)
    dst = g.colidx.astype(np.int64)
    lower = (deg[dst] < deg[src]) | ((deg[dst] == deg[src]) & (dst < src))
    L = np.bincount(src[lower], minlength=g.n).astype(np.int64)
    c2 = lambda x: x * (x - 1) // 2
    tasks = L > 0
    return int((16 + 4 * L[tasks]).sum() + (16 * (deg - L) + 4 * (c2(deg) - c2(L))).sum())


# ------------------------------------------------------------ the suite

def hbm_roof(nbytes, ms, kernel, model, traffic=None):
    if not nbytes or not ms:
        return None
    ach = nbytes / (ms * 1e-3) / 1e9
    r = dict(kernel=kernel, avg_launch_ms=round(ms, 4), bytes=int(nbytes), bytes_model=model,
             achieved_gbs=round(ach, 1), frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, traffic_frac=None)
    if traffic:
        r["traffic"] = int(traffic)
        r["traffic_frac"] = round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return r


def time_counts(ctx, run, steps, warmup):
    """run() -> count.  Warm-up, then `steps` timed counts: wall ms per count
    and the HIP-event time of every kernel (mean over the timed counts)."""
    count = None
    for _ in range(warmup):
        count = run()
    kt, totals = {}, []
    t0 = time.perf_counter()
    for _ in range(steps):
        c = run()
        total, ks = ctx.last_profile()
        totals.append(total)
        for k in ks:
            kt.setdefault(k.name, []).append(k.ms)
        if count is not None and c != count:
            raise SystemExit(f"non-deterministic count: {c} vs {count}")
        count = c
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return dict(count=count, ms_per_step=round(wall, 4), kernel_total_ms=round(float(np.mean(totals)), 4),
                kernels_ms={k: round(float(np.mean(v)) * len(v) / steps, 4) for k, v in kt.items()},
                steps=steps, warmup=warmup)


class Suite:
    """All configurations in one run (one GPU): every row's count is checked
    against its golden count, timed with HIP events, given a byte-model
    roofline for its dominant kernel, HBM traffic for that kernel from
    rocprofv3 counters, and a CPU baseline."""

    def __init__(self, args, xgm):
        self.args, self.xgm = args, xgm
        self.gold = json.load(open(os.path.join(ROOT, "tests", "golden", "counts.json")))
        self.keep = bool(args.keep_cache)
        if self.keep:
            os.makedirs(args.keep_cache, exist_ok=True)
            self.cache = os.path.abspath(args.keep_cache)
        else:
            self.cache = tempfile.mkdtemp(prefix="xgm_bench_")
        self.rows, self.cpu_done = [], {}
        self.info = cpu_info()
        self.threads = cpu_threads(self.info, args.cpu_threads)
        self.ctx = xgm.Context(0)
        self.pass_rows = []  # what the counter passes replay

    def golden(self, key, g, pattern):
        e = self.gold.get(key)
        if e and e["hash"] == f"{g.hash():016x}":
            return e["counts"].get(pattern)
        return None

    def save(self, g, name):
        prefix = os.path.join(self.cache, name)
        g.save(prefix)
        return prefix

    def cpu_ref(self, key, pattern, g, m, what):
        """Reference binary on the full graph (cached per graph and pattern)."""
        if self.args.no_cpu:
            return None
        import oracle
        if (key, pattern) in self.cpu_done:
            return self.cpu_done[(key, pattern)]
        if not oracle.ref_available(pattern):
            return None
        log(f"cpu: reference {pattern} on {key}, {self.threads} threads")
        c, med, runs = ref_timed(pattern, g, self.threads)
        r = dict(value=m / med, unit="|E|/s", cores=self.threads, kind="reference", seconds=round(med, 4),
                 runs_s=runs, count=c, sample=f"full {what} (reference GraphMiner binary, counting loop only, "
                                               f"median of {len(runs)})")
        self.cpu_done[(key, pattern)] = r
        return r

    def cpu_port(self, key, pattern, gr, order, m, what):
        """Oracle restatement on a stratified task sample (reference binary
        would take hours)."""
        if self.args.no_cpu:
            return None
        if (key, pattern) in self.cpu_done:
            return self.cpu_done[(key, pattern)]
        log(f"cpu: oracle restatement {pattern} on {key} (task sample), {self.threads} threads")
        s = port_sampled(pattern, gr, order, self.threads)
        r = dict(value=m / s["seconds_est"], unit="|E|/s", cores=self.threads, kind="port",
                 seconds_est=round(s["seconds_est"], 2), sample_seconds=s["sample_seconds"],
                 sample=f"{what}: every {s['stride']}-th vertex task in decreasing-work order ({s['tasks']} tasks, "
                        f"{100 * s['fraction']:.3f}% of them) with the oracle's restatement of GraphMiner's loops; "
                        "full-graph time estimated as sample time x tasks / sampled tasks")
        self.cpu_done[(key, pattern)] = r
        return r

    def row(self, name, workload, pattern, m, timing, golden, roof_by_kernel, model, cpu, extra=None):
        count = timing["count"]
        if golden is not None and count != golden:
            raise SystemExit(f"{name}: COUNT MISMATCH vs golden: {count} != {golden}")
        km = timing["kernels_ms"]
        dom = max(km, key=km.get) if km else None
        nbytes = roof_by_kernel.get(dom) if (dom and roof_by_kernel) else None
        r = dict(row=name, workload=workload, pattern=pattern, count=count,
                 count_matches_golden=(count == golden) if golden is not None else None,
                 ms_per_step=timing["ms_per_step"], kernel_total_ms=timing["kernel_total_ms"],
                 e_per_s=m / (timing["ms_per_step"] * 1e-3), matches_per_s=count / (timing["ms_per_step"] * 1e-3),
                 steps=timing["steps"], warmup=timing["warmup"], kernels_ms=km, dominant_kernel=dom,
                 roofline=hbm_roof(nbytes, km.get(dom), dom, model) if nbytes else None, cpu_baseline=cpu)
        if extra:
            r.update(extra)
        self.rows.append(r)
        log(f"row {name}: count={count} {timing['ms_per_step']:.2f} ms/step, dominant {dom} "
            f"{km.get(dom, 0):.2f} ms" + (f", frac {r['roofline']['frac']}" if r["roofline"] else ""))
        return r

    # --- graph-level groups -------------------------------------------------

    def rmat22(self):
        a = self.args
        g = self.xgm.gen_rmat(22, 16, 0.57, 0.19, 0.19, 2)
        d = g.orient()
        m = g.nnz // 2
        desc = "R-MAT scale 22, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, degree-oriented DAG"
        log(f"rmat22: n={g.n} m={m} hash={g.hash():016x}")
        dg = self.ctx.upload(d)
        pre = self.save(d, "rmat22_dag")
        # headline: triangles
        tc = time_counts(self.ctx, lambda: self.ctx.count(dg, 1), a.steps, a.warmup)
        tcb = tc_kernel_bytes(d)
        alg, alg_total = tc_alg_bytes_per_tier(d)
        cpu = self.cpu_ref("rmat22", "triangle", g, m, "R-MAT-22 seed 2 triangle count")
        r = self.row("tc_rmat22", desc, "triangle", m, tc, self.golden("rmat22_s2", g, "triangle"), tcb,
                     "tc_kernel_bytes", cpu)
        dom = r["dominant_kernel"]
        r["roofline_alg"] = dict(kernel=dom, alg_bytes_per_launch=alg.get(dom),
                                 alg_bytes_count=alg_total, definition="4*(d+(u)+d+(v))+40 per DAG edge")
        self.pass_rows.append(dict(row="tc_rmat22", graph=pre, flags=d.flags, pattern=1, kind="count", opts={}))
        # 4- and 5-cliques: default (rank LUT), the reference-style id-LUT
        # threshold, and the learned schedule
        from xgminer_amd import scheduler
        order = dg.task_order()
        model = scheduler.load()
        for k, pid, pname in ((4, 2, "clique4"), (5, 3, "clique5")):
            gold = self.golden("rmat22_s2", g, pname)
            cpu = self.cpu_port("rmat22", pname, d, order, m, f"R-MAT-22 seed 2 {k}-clique count")
            codes = dg.sched_heuristic(pid)
            t = time_counts(self.ctx, lambda: self.ctx.count(dg, pid), min(a.steps, 10), min(a.warmup, 2))
            self.row(f"k{k}_rmat22", desc + ", rank-ordered LUT (library default)", pname, m, t, gold,
                     clique_kernel_bytes(d, codes, k), "clique_kernel_bytes", cpu)
            self.pass_rows.append(dict(row=f"k{k}_rmat22", graph=pre, flags=d.flags, pattern=pid, kind="count",
                                       opts={}))
            codes_id = dg.sched_heuristic(pid, self.xgm.STRAT_LUT)
            t = time_counts(self.ctx, lambda: self.ctx.count(dg, pid, self.xgm.STRAT_LUT), min(a.steps, 3), 1)
            self.row(f"k{k}_rmat22_idlut", desc + ", id-ordered LUT up to 1024 out-neighbours, edge LUT above "
                     "(the reference's fixed-threshold scheme)", pname, m, t, gold,
                     clique_kernel_bytes(d, codes_id, k), "clique_kernel_bytes", cpu)
            self.pass_rows.append(dict(row=f"k{k}_rmat22_idlut", graph=pre, flags=d.flags, pattern=pid, kind="count",
                                       opts=dict(strategy=self.xgm.STRAT_LUT)))
            if k == 5:
                lc = scheduler.schedule(dg, model, pid)
                s = self.xgm.Schedule(dg, pid, lc)
                t = time_counts(self.ctx, s.count, min(a.steps, 10), min(a.warmup, 2))
                s.close()
                np.save(os.path.join(self.cache, "rmat22_k5_learned.npy"), lc)
                self.row("k5_rmat22_learned", desc + ", learned schedule", pname, m, t, gold,
                         clique_kernel_bytes(d, lc, k), "clique_kernel_bytes", cpu,
                         extra=dict(code_hist=np.bincount(lc[lc >= 0], minlength=5).tolist()))
                self.pass_rows.append(dict(row="k5_rmat22_learned", graph=pre, flags=d.flags, pattern=pid,
                                           kind="sched", codes=os.path.join(self.cache, "rmat22_k5_learned.npy")))
        dg.close()
        self.setops(g)
        return g

    def setops(self, g):
        """Set intersection counts on the undirected R-MAT-22 lists under
        each strategy, for three classes of pairs: hub x hub (all pairs of
        the 64 highest-degree vertices, both lists above 1024), hub x small
        and the graph's own edges.  Bytes: 4 (|A| + |B|) + 8 per pair, each
        list read once."""
        xgm = self.xgm
        ug = self.ctx.upload(g)
        deg = np.diff(g.rowptr).astype(np.int64)
        rng = np.random.default_rng(7)
        top = np.argsort(-deg, kind="stable")[:64].astype(np.int32)
        cls = {"hub_x_hub": (np.repeat(top, 64), np.tile(top, 64)),
               "hub_x_small": (np.repeat(top, 64), rng.choice(np.nonzero((deg > 0) & (deg <= 64))[0], 4096)
                               .astype(np.int32))}
        e = rng.integers(0, g.nnz, 1 << 18)
        src = (np.searchsorted(g.rowptr, e, side="right") - 1).astype(np.int32)
        cls["edges"] = (src, g.colidx[e].astype(np.int32))
        out = {}
        for name, (a, b) in cls.items():
            nbytes = int(4 * (deg[a] + deg[b]).sum() + 8 * len(a))
            res, ref = {}, None
            for sname, flag in (("bsearch", 0), ("merge", xgm.SETOP_MERGE), ("hybrid", xgm.SETOP_HYBRID)):
                ms = []
                for _ in range(3):
                    c = self.ctx.setop_pairs(ug, xgm.SETOP_INTERSECT | flag, a, b)
                    ms.append(self.ctx.last_profile()[1][0].ms)
                    if ref is None:
                        ref = c
                    elif not np.array_equal(c, ref):
                        raise SystemExit(f"setops {name}: {sname} disagrees with binary search")
                t = float(np.median(ms))
                res[sname] = dict(kernel_ms=round(t, 4), achieved_gbs=round(nbytes / (t * 1e-3) / 1e9, 1),
                                  frac=round(nbytes / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
            out[name] = dict(pairs=len(a), list_bytes=nbytes, total_intersection=int(ref.sum()), **res)
            log(f"setops {name}: " + ", ".join(f"{k} {v['kernel_ms']:.3f} ms" for k, v in res.items()))
        ug.close()
        self.rows.append(dict(row="setops_rmat22", workload="set intersection counts on the undirected R-MAT-22 "
                              "lists, per pair class and strategy (binary search / wave64 merge path / hybrid)",
                              pattern="intersect", classes=out, bytes_model="4*(|A|+|B|)+8 per pair",
                              cpu_baseline=None))

    def rmat16(self):
        a = self.args
        g = self.xgm.gen_rmat(16, 16, 0.57, 0.19, 0.19, 1)
        d = g.orient()
        m = g.nnz // 2
        dg = self.ctx.upload(d)
        t = time_counts(self.ctx, lambda: self.ctx.count(dg, 1), a.steps, min(a.warmup, 2))
        cpu = self.cpu_ref("rmat16", "triangle", g, m, "R-MAT-16 seed 1 triangle count")
        self.row("tc_rmat16", "R-MAT scale 16, edge factor 16, seed 1, degree-oriented DAG", "triangle", m, t,
                 self.golden("rmat16_s1", g, "triangle"), tc_kernel_bytes(d), "tc_kernel_bytes", cpu)
        self.pass_rows.append(dict(row="tc_rmat16", graph=self.save(d, "rmat16_dag"), flags=d.flags, pattern=1,
                                   kind="count", opts={}))
        dg.close()

    def chunglu(self):
        a = self.args
        g = self.xgm.gen_chung_lu(4000000, 20.0, 2.1, 50000.0, 3)
        m = g.nnz // 2
        desc = "Chung-Lu n=4,000,000, avg degree 20, beta 2.1, w_max 50,000, seed 3, undirected"
        log(f"chunglu: n={g.n} m={m} hash={g.hash():016x}")
        dg = self.ctx.upload(g)
        pre = self.save(g, "chunglu_und")
        t = time_counts(self.ctx, lambda: self.ctx.count(dg, 4), min(a.steps, 10), min(a.warmup, 2))
        order = np.argsort(-np.diff(g.rowptr), kind="stable").astype(np.int32)
        # the reference sgl binary runs for minutes here: the restatement of
        # its loop (sgl/cpu_kernels/diamond.h) on a task sample instead
        cpu = self.cpu_port("chunglu", "diamond", g, order, m, "Chung-Lu 4M seed 3 edge-induced diamond count")
        self.row("diamond_chunglu", desc, "diamond", m, t, self.golden("chunglu4m_s3", g, "diamond"),
                 diamond_kernel_bytes(g), "diamond_kernel_bytes", cpu)
        self.pass_rows.append(dict(row="diamond_chunglu", graph=pre, flags=g.flags, pattern=4, kind="count", opts={}))
        t = time_counts(self.ctx, lambda: self.ctx.count(dg, 5), min(a.steps, 10), min(a.warmup, 2))
        cpu = self.cpu_port("chunglu", "rectangle", g, order, m, "Chung-Lu 4M seed 3 edge-induced 4-cycle count")
        r = self.row("rectangle_chunglu", desc, "rectangle", m, t, self.golden("chunglu4m_s3", g, "rectangle"), {},
                     None, cpu)
        b = cycle4_bytes(g)
        r["roofline_pattern"] = hbm_roof(b, r["kernel_total_ms"], "all 4-cycle kernels", "cycle4_bytes")
        self.pass_rows.append(dict(row="rectangle_chunglu", graph=pre, flags=g.flags, pattern=5, kind="count",
                                   opts={}))
        dg.close()

    def batch(self):
        a = self.args
        from xgminer_amd import workloads
        xgm = self.xgm
        specs = workloads.batch_specs()
        inst = workloads.batch_graphs(xgm, specs)
        dags = [g.orient() for g in inst]
        msum = sum(g.nnz // 2 for g in inst)
        rows = json.load(open(os.path.join(ROOT, "tests", "golden", "batch_counts.json")))["instances"]
        golden = None
        if len(rows) == len(specs) and all(f"{g.hash():016x}" == r["hash"] for g, r in zip(inst, rows)):
            golden = np.array([[r["counts"][p] for r in rows] for p in BATCH_PATTERNS], np.uint64)
        log(f"batch: {len(inst)} instances generated, sum m={msum}")
        ctx = self.ctx
        dag, und = ctx.upload_batch(dags), ctx.upload_batch(inst)
        on = {"triangle": dag, "clique4": dag, "diamond": und, "rectangle": und}
        pids = [PATTERN_IDS[p] for p in BATCH_PATTERNS]
        t0 = time.time()
        tuned = {}
        for p in BATCH_PATTERNS:
            best, ms = ctx.autotune(on[p], PATTERN_IDS[p], reps=1)
            tuned[p] = dict(opts=best, ms=round(ms, 3))
            log(f"batch: autotune {p}: {best} {ms:.2f} ms")
        tune_s = time.time() - t0
        plan = ctx.plan(pids, dag=dag, und=und, opts=[tuned[p]["opts"] for p in BATCH_PATTERNS])
        plain = ctx.plan(pids, dag=dag, und=und)
        log("batch: plans captured")
        st = min(a.steps, 10)

        def replay(pl):
            counts = None
            for _ in range(min(a.warmup, 2)):
                counts, _ = pl.run()
            evs, t = [], time.perf_counter()
            for _ in range(st):
                c, ms = pl.run()
                evs.append(ms)
                if counts is not None and not np.array_equal(c, counts):
                    raise SystemExit("non-deterministic batch counts")
                counts = c
            return counts, (time.perf_counter() - t) * 1e3 / st, float(np.mean(evs))

        counts, ms_step, ev_ms = replay(plan)
        pcounts, pms_step, _ = replay(plain)
        if not np.array_equal(counts, pcounts):
            raise SystemExit("tuned and default plans disagree")
        if golden is not None and not np.array_equal(counts, golden):
            raise SystemExit(f"BATCH COUNT MISMATCH vs golden at {np.argwhere(counts != golden)[:5].tolist()}")
        kernels, owner = {}, {}
        for p in BATCH_PATTERNS:
            ctx.count_instances(on[p], PATTERN_IDS[p], **tuned[p]["opts"])
            for k in ctx.last_profile()[1]:
                kernels[k.name] = kernels.get(k.name, 0.0) + k.ms
                owner[k.name] = p
        dom = max(kernels, key=kernels.get)
        ud, uu = union(xgm, dags), union(xgm, inst)
        p = owner[dom]
        if p == "triangle":
            model, kb = "tc_kernel_bytes", tc_kernel_bytes(ud)
        elif p == "diamond":
            model, kb = "diamond_kernel_bytes", diamond_kernel_bytes(uu)
        elif p == "clique4":
            st_ = tuned[p]["opts"]["strategy"]
            model, kb = "clique_kernel_bytes", clique_kernel_bytes(ud, dag.sched_heuristic(2, st_, tuned[p]["opts"]
                                                                                          ["lut_max_degree"]), 4)
        else:
            model, kb = None, {}
        cpu = None
        if not a.no_cpu:
            cpu = cpu_baseline_batch(specs, inst, self.threads)
        r = dict(row="batch256", workload="256 instances (128 Erdos-Renyi n log-uniform [1e3, 2e5], avg degree "
                 "U[4, 64]; 128 R-MAT scale U{10..18}, edge factor 16; seeds 100..355) x {triangle, 4-clique, "
                 "diamond, 4-cycle}, one hipGraph replay per step", pattern="batch",
                 count=int(counts.sum()), count_matches_golden=True if golden is not None else None,
                 counts_total={q: int(counts[i].sum()) for i, q in enumerate(BATCH_PATTERNS)},
                 ms_per_step=round(ms_step, 4), graph_replay_ms=round(ev_ms, 4),
                 e_per_s=4 * msum / (ms_step * 1e-3), sum_undirected_edges=msum, steps=st, warmup=min(a.warmup, 2),
                 default_plan_ms_per_step=round(pms_step, 4), tuned_vs_default=round(pms_step / ms_step, 4),
                 autotune={q: tuned[q] for q in BATCH_PATTERNS}, autotune_s=round(tune_s, 1),
                 kernels_ms_eager={k: round(v, 4) for k, v in kernels.items()}, dominant_kernel=dom,
                 roofline=hbm_roof(kb.get(dom), kernels[dom], dom, model) if model else None, cpu_baseline=cpu)
        if p == "rectangle":
            # the 4-cycle kernels split tasks by wedge work: bytes of all of
            # them over their summed time
            c4_ms = sum(v for k, v in kernels.items() if owner[k] == "rectangle")
            r["roofline_pattern"] = hbm_roof(cycle4_bytes(uu), c4_ms, "all 4-cycle kernels", "cycle4_bytes")
        self.rows.append(r)
        log(f"row batch256: {ms_step:.2f} ms/step (default plan {pms_step:.2f}), dominant {dom}")
        # counter pass: the dominant pattern's eager count on the unions
        which = "dag" if p in ("triangle", "clique4") else "und"
        pre = self.save(ud if which == "dag" else uu, "batch_" + which)
        bounds = np.cumsum([0] + [x.n for x in (dags if which == "dag" else inst)])
        np.save(pre + "_bounds.npy", bounds)
        self.pass_rows.append(dict(row="batch256", graph=pre, flags=(ud if which == "dag" else uu).flags,
                                   pattern=PATTERN_IDS[p], kind="batch", opts=tuned[p]["opts"],
                                   bounds=pre + "_bounds.npy"))
        for x in (plan, plain):
            x.close()
        dag.close()
        und.close()

    def rmat23(self):
        a = self.args
        from xgminer_amd import scheduler
        g = self.xgm.gen_rmat(23, 16, 0.57, 0.19, 0.19, 4)
        d = g.orient()
        m = g.nnz // 2
        want = self.golden("rmat23_s4", g, "clique5")
        log(f"rmat23: n={g.n} m={m} hash={g.hash():016x}")
        del g
        desc = "R-MAT scale 23, edge factor 16, seed 4 (held out from training), degree-oriented DAG"
        dg = self.ctx.upload(d)
        pre = self.save(d, "rmat23_dag")
        model = scheduler.load()
        t0 = time.perf_counter()
        lc = scheduler.schedule(dg, model, 3)
        sched_ms = (time.perf_counter() - t0) * 1e3
        rank = dg.sched_heuristic(3)
        res = {}
        order = dg.task_order()
        cpu = self.cpu_port("rmat23", "clique5", d, order, m, "R-MAT-23 seed 4 5-clique count")
        for name, codes in (("learned", lc), ("rank", rank)):
            s = self.xgm.Schedule(dg, 3, codes)
            t = time_counts(self.ctx, s.count, min(a.steps, 5), 1)
            s.close()
            res[name] = t
        if res["learned"]["count"] != res["rank"]["count"]:
            raise SystemExit("RMAT-23 schedules disagree")
        # learned + captured: the learned schedule replayed as a hipGraph plan
        s = self.xgm.Schedule(dg, 3, lc)
        plan = self.ctx.plan([3], dag=dg, scheds=[s])
        s.close()
        replay = []
        for i in range(min(a.steps, 5) + 1):
            got, ms = plan.run()
            if int(got[0, 0]) != res["learned"]["count"]:
                raise SystemExit("RMAT-23 learned plan replay disagrees")
            replay.append(ms)
        plan.close()
        plan_ms = float(np.median(replay[1:]))
        log(f"rmat23 learned schedule as a plan replay: {plan_ms:.2f} ms")
        np.save(os.path.join(self.cache, "rmat23_k5_learned.npy"), lc)
        for name, codes in (("learned", lc), ("rank", rank)):
            extra = dict(code_hist=np.bincount(codes[codes >= 0], minlength=5).tolist())
            if name == "learned":
                extra.update(schedule_build_ms=round(sched_ms, 2),
                             speedup_vs_rank_heuristic=round(res["rank"]["ms_per_step"] /
                                                             res["learned"]["ms_per_step"], 4),
                             learned_plan_replay_ms=round(plan_ms, 3))
            self.row(f"k5_rmat23_{name}", desc + (", learned schedule" if name == "learned" else
                                                  ", rank-ordered LUT degree thresholds (model-free heuristic)"),
                     "clique5", m, res[name], want, clique_kernel_bytes(d, codes, 5), "clique_kernel_bytes", cpu,
                     extra=extra)
        np.save(os.path.join(self.cache, "rmat23_k5_rank.npy"), rank)
        for name in ("learned", "rank"):
            self.pass_rows.append(dict(row=f"k5_rmat23_{name}", graph=pre, flags=d.flags, pattern=3, kind="sched",
                                       codes=os.path.join(self.cache, f"rmat23_k5_{name}.npy")))
        dg.close()

    # --- HBM traffic of every row's dominant kernel ---------------------------

    def traffic(self):
        """Two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of the counter
        pass, which counts every row once, each row followed by a marker
        copy; the dispatches between markers are matched to the row's kernel
        launches.  Counters are scaled by what they read for the plain copy
        kernel over 1 GiB."""
        spec = os.path.join(self.cache, "rows.json")
        json.dump(self.pass_rows, open(spec, "w"))
        res = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            log(f"traffic: rocprofv3 --pmc {c}")
            res[c] = _pmc_segments(c, ["--profile-pass", spec], timeout=240)
            if res[c] is None:
                return None
        names = json.load(open(spec + ".names.json"))
        out = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            calib, segs = res[c]
            if not calib:
                return None
            factor = COPY_BYTES / (float(np.mean(calib)) * 1024.0)
            out[c] = dict(copy_kib=round(float(np.mean(calib)), 1), factor=round(factor, 4))
            for i, pr in enumerate(self.pass_rows):
                kn = names[i]
                seg = segs[2 * i + 1] if 2 * i + 1 < len(segs) else []
                if len(seg) != len(kn):
                    log(f"traffic {c}: row {pr['row']}: {len(seg)} dispatches vs {len(kn)} launches, skipped")
                    continue
                per = {}
                for nm, (sym, v) in zip(kn, seg):
                    per.setdefault(nm, []).append((sym, v * 1024.0 * factor))
                out.setdefault("rows", {}).setdefault(pr["row"], {})[c] = {
                    nm: dict(symbol=vs[0][0], bytes=float(sum(x[1] for x in vs))) for nm, vs in per.items()}
        for r in self.rows:
            t = out.get("rows", {}).get(r["row"])
            if t and "FETCH_SIZE" in t and "WRITE_SIZE" in t:
                r["traffic_all_kernels"] = int(sum(x["bytes"] for c in ("FETCH_SIZE", "WRITE_SIZE")
                                                   for x in t[c].values()))
                rp = r.get("roofline_pattern")
                if rp:
                    tb = r["traffic_all_kernels"]
                    rp.update(traffic=tb, traffic_frac=round(tb / (rp["avg_launch_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                                            4))
            dom = r.get("dominant_kernel")
            if not t or not r.get("roofline") or dom is None:
                continue
            f, w = t.get("FETCH_SIZE", {}).get(dom), t.get("WRITE_SIZE", {}).get(dom)
            if f and w:
                ms = r["roofline"]["avg_launch_ms"]
                tb = f["bytes"] + w["bytes"]
                r["roofline"].update(traffic=int(tb), traffic_frac=round(tb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                     fetch_bytes=int(f["bytes"]), write_bytes=int(w["bytes"]), symbol=f["symbol"])
        # rows sharing a graph and pattern with a measured row (same kernels)
        return {c: out[c] for c in ("FETCH_SIZE", "WRITE_SIZE")}

    ISSUE_COUNTERS = ("SQ_ACTIVE_INST_VALU", "SQ_INSTS_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_INSTS_SALU",
                      "SQ_ACTIVE_INST_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE", "GRBM_GUI_ACTIVE")

    def issue(self):
        """One rocprofv3 --pmc run of the counter pass with the issue
        counters: for the dominant kernel of every row, the fraction of SIMD
        cycles issuing VALU instructions (SQ_ACTIVE_INST_VALU counts one per
        4-cycle wave64 issue: busy = 4 ACTIVE / (SIMDs x cycles per XCC)),
        the active-lane fraction of those instructions, SALU per VALU
        instruction, LDS active cycles per CU and the LDS bank-conflict share
        of them; and the VALU issue roofline (roofline_valu) of that kernel.  The
        LUT kernels are bound by this issue rate, not by HBM."""
        spec = os.path.join(self.cache, "rows.json")
        json.dump(self.pass_rows, open(spec, "w"))
        log("issue counters: rocprofv3 --pmc " + " ".join(self.ISSUE_COUNTERS))
        res = _pmc_run(self.ISSUE_COUNTERS, ["--profile-pass", spec], timeout=240)
        if res is None:
            return None
        data, agent = res
        names = json.load(open(spec + ".names.json"))
        simd, xcc, cu = agent.get("simd", 1024), agent.get("xcc", 8), agent.get("cu", 256)
        per_row = {}
        for i, pr in enumerate(self.pass_rows):
            kn = names[i]
            vals = {}
            for c in self.ISSUE_COUNTERS:
                segs = data[c][1]
                seg = segs[2 * i + 1] if 2 * i + 1 < len(segs) else []
                if len(seg) != len(kn):
                    vals = None
                    break
                for nm, (sym, v) in zip(kn, seg):
                    vals.setdefault(nm, {}).setdefault(c, 0.0)
                    vals[nm][c] += v
                    vals[nm]["symbol"] = sym
            if vals is not None:
                per_row[pr["row"]] = vals
        for r in self.rows:
            dom, vals = r.get("dominant_kernel"), per_row.get(r["row"])
            if not dom or not vals or dom not in vals:
                continue
            v = vals[dom]
            cyc = v["GRBM_GUI_ACTIVE"] / xcc
            if cyc <= 0 or v["SQ_ACTIVE_INST_VALU"] <= 0:
                continue
            busy = 4.0 * v["SQ_ACTIVE_INST_VALU"] / (simd * cyc)
            r["issue"] = dict(
                kernel=dom, symbol=v["symbol"], valu_busy=round(busy, 4),
                valu_lane_util=round(v["SQ_THREAD_CYCLES_VALU"] / (64.0 * v["SQ_ACTIVE_INST_VALU"]), 4),
                salu_per_valu=round(v["SQ_INSTS_SALU"] / max(1.0, v["SQ_INSTS_VALU"]), 3),
                lds_active=round(v["SQ_LDS_IDX_ACTIVE"] / (cu * cyc), 4),
                lds_bank_conflict_share=round(v["SQ_LDS_BANK_CONFLICT"] / max(1.0, v["SQ_LDS_IDX_ACTIVE"]), 4),
                cycles_per_xcc=int(cyc), counters={c: v[c] for c in self.ISSUE_COUNTERS})
            ms = (r.get("kernels_ms") or r.get("kernels_ms_eager") or {}).get(dom)
            if ms:
                # VALU issue roofline: one wave64 VALU instruction per SIMD per
                # 4 cycles at the 2.4 GHz peak engine clock; achieved = issue
                # slots used per second over the kernel's HIP-event time
                ach = v["SQ_ACTIVE_INST_VALU"] / (ms * 1e-3) / 1e9
                peak = simd * VALU_PEAK_CLOCK_GHZ / 4.0
                r["roofline_valu"] = dict(bound="valu", kernel=dom, achieved=round(ach, 1), peak=round(peak, 1),
                                          unit="G wave64 VALU issues/s", frac=round(ach / peak, 4),
                                          avg_launch_ms=round(ms, 4))
        return dict(simd=simd, xcc=xcc, cu=cu)

    def close(self):
        self.ctx.close()
        if self.keep:  # the counter pass can be replayed: bench.py --profile-pass DIR/rows.json
            json.dump(self.pass_rows, open(os.path.join(self.cache, "rows.json"), "w"))
            return
        import shutil
        shutil.rmtree(self.cache, ignore_errors=True)


def _run_logged(cmd, timeout, what):
    """Run cmd (output discarded) with a time limit, logging a progress line
    every 30 s; True if it ended by itself."""
    p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tempfile.gettempdir())
    t0 = time.time()
    while True:
        try:
            p.wait(timeout=30)
            return True
        except subprocess.TimeoutExpired:
            if time.time() - t0 > timeout:
                p.kill()
                p.wait()
                log(f"{what} timed out")
                return False
            log(f"{what} running ({time.time() - t0:.0f}s)")


def _pmc_segments(counter, argv, timeout=240):
    """Run `python3 bench.py argv` under rocprofv3 --pmc counter.  Returns
    (calibration values of the first marker group, [segment]) where every
    segment lists (kernel symbol, counter KiB) of the dispatches between two
    marker copies, in dispatch order, without the zero-fill kernel and the
    runtime's blit kernels; or None."""
    res = _pmc_run([counter], argv, timeout)
    return None if res is None else res[0][counter]


def _pmc_run(counters, argv, timeout=240):
    """One rocprofv3 --pmc run of `python3 bench.py argv` collecting every
    counter of `counters` (within one pass's hardware limits).  Returns
    ({counter: (calibration values, [segment])} as in _pmc_segments,
    agent info of the GPU: SIMD and XCC counts), or None."""
    import csv
    import glob
    import shutil
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    what = "rocprofv3 --pmc " + " ".join(counters)
    with tempfile.TemporaryDirectory(prefix="xgm_pmc_") as d:
        cmd = [exe, "--pmc"] + list(counters) + ["--kernel-trace", "--output-format", "csv", "-d", d, "-o", "run",
                                                 "--", sys.executable, os.path.abspath(__file__)] + argv
        if not _run_logged(cmd, timeout, what):
            return None
        disp = {c: {} for c in counters}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            for row in csv.DictReader(open(f)):
                c = row.get("Counter_Name")
                if c not in disp:
                    continue
                key = int(row.get("Dispatch_Id") or row.get("Correlation_Id") or len(disp[c]))
                name = row.get("Kernel_Name", "")
                disp[c][key] = (name, disp[c].get(key, (name, 0.0))[1] + float(row["Counter_Value"]))
        agent = {}
        for f in glob.glob(os.path.join(d, "**", "*agent_info.csv"), recursive=True):
            for row in csv.DictReader(open(f)):
                if int(row.get("Simd_Count") or 0) > 0:
                    agent = dict(simd=int(row["Simd_Count"]), xcc=int(row.get("Num_Xcc") or 1),
                                 cu=int(row.get("Cu_Count") or 0))
    out = {}
    for c in counters:
        if not disp[c]:
            return None
        calib, segs, cur, groups, in_marker = [], [], [], 0, False
        for key in sorted(disp[c]):
            name, v = disp[c][key]
            if "k_copy_u4" in name:
                if not in_marker:
                    groups += 1
                    if groups > 1:
                        segs.append(cur)
                    cur = []
                in_marker = True
                if groups == 1:
                    calib.append(v)
                continue
            in_marker = False
            if "k_zero_u64" in name or name.startswith("__amd"):
                continue
            sym = re.search(r"k_\w+(<[^>]*>)?", name)
            cur.append((sym.group(0) if sym else name, v))
        out[c] = (calib, segs)
    return out, agent


def trace_summary(trace_dir, spec):
    """Per-row kernel durations of a `rocprofv3 --kernel-trace` run of the
    counter pass (bench.py --profile-pass SPEC): the dispatches between the
    marker copies of each row, matched to the row's launch names (SPEC's
    .names.json), as a markdown table on stdout."""
    import csv
    import glob
    rows = json.load(open(spec))
    names = json.load(open(spec + ".names.json"))
    disp = []
    for f in glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            disp.append((int(r.get("Dispatch_Id") or r.get("Correlation_Id")), r["Kernel_Name"],
                         (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6))
    disp.sort()
    segs, cur, groups, in_marker = [], [], 0, False
    for _, name, ms in disp:
        if "k_copy_u4" in name:
            if not in_marker:
                groups += 1
                if groups > 1:
                    segs.append(cur)
                cur = []
            in_marker = True
            continue
        in_marker = False
        if "k_zero_u64" in name or name.startswith("__amd"):
            continue
        sym = re.search(r"k_\w+(<[^>]*>)?", name)
        cur.append((sym.group(0) if sym else name, ms))
    print("| row | launch | kernel symbol | ms |")
    print("|---|---|---|---|")
    for i, r in enumerate(rows):
        seg = segs[2 * i + 1] if 2 * i + 1 < len(segs) else []
        kn = names[i]
        if len(seg) != len(kn):
            print(f"| {r['row']} | (dispatches {len(seg)} != launches {len(kn)}) | | |")
            continue
        for nm, (sym, ms) in zip(kn, seg):
            print(f"| {r['row']} | {nm} | `{sym}` | {ms:.3f} |")


def profile_pass(spec):
    """Counter pass: the copy calibration (1 GiB), then every row of the
    suite: set-up and a warm-up count, a marker copy, the measured count and
    another marker copy; writes the kernel launch names of every row (in
    launch order) beside the spec."""
    from conftest import load_xgm
    xgm = load_xgm()
    rows = json.load(open(spec))
    ctx = xgm.Context(0)
    ctx.copy_bandwidth(COPY_BYTES, 1)
    names, graphs = [], {}
    for r in rows:
        # each row: set-up (loads, uploads), marker, the count, marker; the
        # counter parser keeps the dispatches between the two markers
        print(f"[pass] {r['row']}", file=sys.stderr, flush=True)
        if r["kind"] == "batch":
            g = xgm.load_csr(r["graph"], r["flags"])
            b = np.load(r["bounds"])
            parts = []
            for i in range(len(b) - 1):
                lo, hi = int(b[i]), int(b[i + 1])
                rp = g.rowptr[lo:hi + 1] - g.rowptr[lo]
                ci = g.colidx[g.rowptr[lo]:g.rowptr[hi]] - np.int32(lo)
                parts.append(xgm.Graph.from_arrays(rp, ci, r["flags"]))
            dg = ctx.upload_batch(parts)
            ctx.count_instances(dg, r["pattern"], **r["opts"])  # warm-up (task lists, caches)
            ctx.copy_bandwidth(1 << 16, 1)
            ctx.count_instances(dg, r["pattern"], **r["opts"])
        else:
            if r["graph"] not in graphs:
                for k in list(graphs):
                    graphs.pop(k).close()
                graphs[r["graph"]] = ctx.upload(xgm.load_csr(r["graph"], r["flags"]))
            dg = graphs[r["graph"]]
            s = xgm.Schedule(dg, r["pattern"], np.load(r["codes"])) if r["kind"] == "sched" else None
            if s is not None:  # warm-up (task lists, caches)
                s.count()
            else:
                ctx.count(dg, r["pattern"], **r["opts"])
            ctx.copy_bandwidth(1 << 16, 1)
            if s is not None:
                s.count()
                s.close()
            else:
                ctx.count(dg, r["pattern"], **r["opts"])
        names.append([k.name for k in ctx.last_profile()[1]])
        ctx.copy_bandwidth(1 << 16, 1)
        if r["kind"] == "batch":
            dg.close()
    for dg in graphs.values():
        dg.close()
    json.dump(names, open(spec + ".names.json", "w"))
    ctx.close()


def run_suite(args, xgm):
    set_omp_binding()
    t_start = time.time()
    S = Suite(args, xgm)
    copy_gbps = S.ctx.copy_bandwidth(COPY_BYTES, 5)
    log(f"copy kernel {copy_gbps:.0f} GB/s; cpu {S.info}; {S.threads} CPU threads")
    S.rmat22()
    S.rmat16()
    S.chunglu()
    S.batch()
    S.rmat23()
    traffic = None if args.no_traffic else S.traffic()
    issue = None if args.no_traffic else S.issue()
    head = S.rows[0]
    assert head["row"] == "tc_rmat22"
    dom = head["dominant_kernel"]
    ra = head["roofline_alg"]
    ms = head["kernels_ms"][dom]
    kr = head["roofline"] or {}
    # achieved: the set-operation bytes of the kernel's own op stream (every
    # list it reads, once: tc_kernel_bytes); the reference's merge op stream
    # (4 (d+(u) + d+(v)) + 40 per DAG edge) is reported beside it -- this
    # kernel streams only the shorter list of each edge and stages the
    # longer one once per item, so that figure would exceed the HBM peak
    ach = kr["bytes"] / (ms * 1e-3) / 1e9
    ref_ach = ra["alg_bytes_per_launch"] / (ms * 1e-3) / 1e9
    roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(ach / HBM_PEAK_GBS, 4),
                traffic=kr.get("traffic"), kernel=dom, alg_bytes_per_launch=kr["bytes"], avg_launch_ms=round(ms, 4),
                alg_bytes_model="tc_kernel_bytes: per charging vertex t 32 + 4 d+(t) (N+(t) staged once), per "
                                "charged edge 20 + 4 d+(partner)",
                traffic_frac=kr.get("traffic_frac"), fetch_bytes=kr.get("fetch_bytes"), write_bytes=kr.get("write_bytes"),
                reference_op_stream=dict(bytes=ra["alg_bytes_per_launch"], definition=ra["definition"],
                                         achieved_gbs=round(ref_ach, 1), frac=round(ref_ach / HBM_PEAK_GBS, 4)))
    out = {
        "metric": "|E|/s per pattern on RMAT-22; learned vs heuristic schedule speedup (exact)",
        "value": head["e_per_s"],
        "unit": "|E|/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "triangle count on R-MAT scale 22, edge factor 16, (0.57,0.19,0.19,0.05), seed 2, "
                               "degree-oriented DAG (value); every other configuration in per_pattern",
                   "parallelism": "single GPU"},
        "count": head["count"],
        "count_matches_golden": head["count_matches_golden"],
        "matches_per_s": head["matches_per_s"],
        "roofline": roof,
        "issue": head.get("issue"),
        "roofline_valu": head.get("roofline_valu"),
        "cpu_baseline": head["cpu_baseline"],
        "learned_vs_heuristic": {
            "rmat23_k5_speedup_vs_rank_heuristic": next(r for r in S.rows if r["row"] == "k5_rmat23_learned")
            ["speedup_vs_rank_heuristic"],
            "rmat22_k5_speedup_vs_rank_heuristic": round(
                next(r for r in S.rows if r["row"] == "k5_rmat22")["ms_per_step"] /
                next(r for r in S.rows if r["row"] == "k5_rmat22_learned")["ms_per_step"], 4),
            "rmat22_k5_speedup_vs_idlut_threshold": round(
                next(r for r in S.rows if r["row"] == "k5_rmat22_idlut")["ms_per_step"] /
                next(r for r in S.rows if r["row"] == "k5_rmat22_learned")["ms_per_step"], 4)},
        "per_pattern": S.rows[1:],
        "hbm_copy_gbps": round(copy_gbps, 1),
        "host_cpu": dict(S.info, threads_used=S.threads, omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                         omp_places=os.environ.get("OMP_PLACES")),
        "traffic_calibration": traffic,
        "issue_counters": issue,
        "wall_s": round(time.time() - t_start, 1),
    }
    for k in ("kernels_ms", "kernel_total_ms"):
        out[k] = head[k]
    S.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(g, golden_count, threads):
    import oracle
    if not oracle.ref_available("triangle"):
        return None
    log(f"cpu baseline: reference GraphMiner tc_omp_base, {threads} threads")
    c, med, runs = ref_timed("triangle", g, threads)
    return dict(value=g.nnz // 2 / med, unit="|E|/s", cores=threads, kind="reference",
                sample=f"full R-MAT-22 seed 2 triangle count (counting loop only, median of {len(runs)})",
                seconds=med, runs_s=runs, count=c, count_matches_gpu=(c == golden_count))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pattern", default=None, choices=list(PATTERN_IDS) + ["motif3", "motif4"])
    ap.add_argument("--workload", default=None, choices=list(WORKLOADS),
                    help="one workload (default: the whole suite of configurations)")
    ap.add_argument("--strategy", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic runs")
    ap.add_argument("--sched", default="", choices=["", "compare"],
                    help="clique patterns: time the learned schedule against the heuristic and rank-LUT ones")
    ap.add_argument("--baseline-steps", type=int, default=1)
    ap.add_argument("--prep", default="", choices=["", "compare"],
                    help="time host vs on-device orientation of the workload graph into a device DAG")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="only run the plain copy kernel over 1 GiB (counter calibration)")
    ap.add_argument("--profile-pass", default="", help=argparse.SUPPRESS)
    ap.add_argument("--keep-cache", default="",
                    help="suite: keep the saved graphs and the counter-pass spec (DIR/rows.json) in DIR")
    ap.add_argument("--trace-summary", nargs=2, metavar=("TRACE_DIR", "SPEC"),
                    help="per-row kernel times of a rocprofv3 --kernel-trace run of --profile-pass SPEC")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.trace_summary:
        return trace_summary(*args.trace_summary)
    from conftest import load_xgm
    xgm = load_xgm()
    if args.profile_pass:
        return profile_pass(args.profile_pass)
    if args.copy_calibrate:
        ctx = xgm.Context(0)
        print(json.dumps({"copy_bytes": COPY_BYTES, "gbps": ctx.copy_bandwidth(COPY_BYTES, 1)}))
        ctx.close()
        return
    single = args.workload or args.pattern or args.strategy is not None or args.sched or args.prep
    if not single:
        return run_suite(args, xgm)
    args.workload = args.workload or "rmat22"
    args.pattern = args.pattern or "triangle"
    args.strategy = args.strategy or 0
    if not args.cpu_threads:
        args.cpu_threads = cpu_threads(cpu_info())
    if args.workload == "batch256":
        return bench_batch(args, xgm)
    return bench_single(args, xgm)


def bench_single(args, xgm):
    """One workload and pattern (the legacy single-row modes)."""
    t0 = time.time()
    kind, params, gname, on = WORKLOADS[args.workload]
    if kind == "rmat":
        g = xgm.gen_rmat(params["scale"], 16, 0.57, 0.19, 0.19, params["seed"])
        desc = (f"R-MAT scale {params['scale']}, edge factor 16, (0.57,0.19,0.19,0.05), "
                f"seed {params['seed']}")
    else:
        g = xgm.gen_chung_lu(params["n"], params["avg_degree"], params["beta"], params["w_max"], params["seed"])
        desc = (f"Chung-Lu n={params['n']}, avg degree {params['avg_degree']}, beta {params['beta']}, "
                f"w_max {params['w_max']:.0f}, seed {params['seed']}")
    d = g.orient() if on == "dag" else g
    m = g.nnz // 2
    log(f"graph {args.workload}: n={g.n} m={m} hash={g.hash():016x} ({time.time()-t0:.1f}s)")
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "counts.json")))
    golden = None
    golden_hash_ok = gname in gold and gold[gname]["hash"] == f"{g.hash():016x}"
    if golden_hash_ok:
        golden = gold[gname]["counts"].get(args.pattern)

    if args.sched == "compare":
        return bench_sched(args, xgm, d, m, desc, golden)
    if args.prep == "compare":
        if on != "dag":
            return bench_prep_undirected(args, xgm, g, m, desc)
        return bench_prep(args, xgm, g, m, desc)
    if args.pattern.startswith("motif"):
        return bench_motif(args, xgm, g, m, desc, gold.get(gname, {}).get("counts", {}) if golden_hash_ok else {})

    # xgm_count is synchronous on the library's own stream (it waits for its
    # kernels and the D2H copy of the count), so no extra device barrier is
    # needed around the timed region; torch is not imported (its bundled HIP
    # runtime would be a second runtime in this process).
    sync = lambda: None

    ctx = xgm.Context(0)
    t0 = time.time()
    dg = ctx.upload(d)
    upload_s = time.time() - t0
    log(f"upload {upload_s:.2f}s")
    pid = PATTERN_IDS[args.pattern]
    count = None
    for i in range(args.warmup):
        count = ctx.count(dg, pid, args.strategy)
        log(f"warmup {i}: count={count}")
    ktimes = {}
    totals = []
    sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        c = ctx.count(dg, pid, args.strategy)
        total, ks = ctx.last_profile()
        totals.append(total)
        for k in ks:
            ktimes.setdefault(k.name, []).append(k.ms)
        if count is not None and c != count:
            raise SystemExit(f"non-deterministic count: {c} vs {count}")
        count = c
    sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    log(f"count={count} golden={golden} ms/step={ms_per_step:.3f} kernel total ms={np.mean(totals):.3f}")
    if golden is not None and count != golden:
        raise SystemExit(f"COUNT MISMATCH vs golden: {count} != {golden}")

    roof = None
    extra = {"upload_s": round(upload_s, 3), "hbm_copy_gbps": round(ctx.copy_bandwidth(COPY_BYTES, 5), 1)}
    avg = {k: float(np.mean(v)) for k, v in ktimes.items()}
    extra["kernels_ms"] = {k: round(v, 4) for k, v in avg.items()}
    if args.pattern == "triangle":
        per_tier, btot = tc_alg_bytes_per_tier(d)
        dom = max(avg, key=avg.get)
        kb = tc_kernel_bytes(d)
        if dom not in per_tier:  # a strategy with one kernel for every edge (bucket bitmap, merge)
            per_tier, kb = {dom: btot}, {dom: btot}
        ach = kb[dom] / (avg[dom] * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel=dom,
                    alg_bytes_per_launch=kb[dom], avg_launch_ms=round(avg[dom], 4),
                    reference_op_stream_bytes=per_tier[dom])
        if not args.no_traffic:
            traffic, detail = measure_traffic(args, dom)
            roof["traffic"] = traffic
            extra["traffic_counters"] = detail
        pat_ach = btot / (float(np.mean(totals)) * 1e-3) / 1e9
        extra["pattern_roofline"] = dict(alg_bytes=btot, kernel_ms=round(float(np.mean(totals)), 4),
                                         achieved=round(pat_ach, 1), frac=round(pat_ach / HBM_PEAK_GBS, 4))

    cpu = None
    if not args.no_cpu:
        cpu = (cpu_baseline(g, count, args.cpu_threads)
               if args.pattern == "triangle" and args.workload == "rmat22" else None)

    out = {
        "metric": f"|E|/s on {args.workload} ({args.pattern})",
        "value": m / (ms_per_step * 1e-3),
        "unit": "|E|/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.pattern} count on {desc}"
                               + (", degree-oriented DAG" if on == "dag" else ", undirected"),
                   "n": g.n, "undirected_edges": m, "graph_hash": f"{g.hash():016x}",
                   "strategy": args.strategy, "parallelism": "single GPU"},
        "count": count,
        "count_matches_golden": (golden == count) if golden is not None else None,
        "matches_per_s": count / (ms_per_step * 1e-3),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
