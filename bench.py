"""Benchmark: |E|/s per pattern on R-MAT scale 22 (edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 2, degree-oriented int32 CSR) on one
MI355X.  A step is one full count of the pattern over the resident graph.

--workload batch256: the 256-instance batch (x-gminer_amd/workloads.py);
a step is one replay of the hipGraph plan counting {triangle, 4-clique,
diamond, 4-cycle} on every instance (1,024 counts) with autotuned options.

Prints one JSON line (rank 0): metric/value/unit, ms_per_step, the roofline
of the dominant kernel (algorithmic bytes / HIP-event duration vs 8 TB/s),
and the reference CPU path timed on this host's cores (cpu_baseline).
Progress goes to stderr."""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
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


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


def cpu_baseline_batch(specs, inst, threads, budget_s=20.0):
    """Reference GraphMiner OpenMP binaries on a bounded sample of the batch:
    instances taken alternately from the ER and RMAT halves, each counted for
    all four patterns, until the counting loops have run budget_s seconds."""
    import oracle
    if not all(oracle.ref_available(p) for p in BATCH_PATTERNS):
        return None
    er = [i for i, s in enumerate(specs) if s[1] == "er"]
    rm = [i for i, s in enumerate(specs) if s[1] == "rmat"]
    order = [x for pair in zip(er, rm) for x in pair]
    tot_t, tot_e, used = 0.0, 0, []
    for i in order:
        g = inst[i]
        for p in BATCH_PATTERNS:
            _, t = oracle.ref_count_graph(p, g, threads=threads, timeout=600)
            tot_t += t
            tot_e += g.nnz // 2
        used.append(specs[i][0])
        log(f"cpu baseline: seed {specs[i][0]} done, {tot_t:.1f}s")
        if tot_t >= budget_s:
            break
    return dict(value=tot_e / tot_t, unit="|E|/s", cores=threads, kind="reference",
                sample=f"{len(used)} batch instances (seeds {used}), all four patterns each, "
                       "GraphMiner tc/clique/sgl OpenMP binaries (counting loops only)",
                seconds=round(tot_t, 3))


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
        try:
            subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=timeout,
                           cwd=tempfile.gettempdir())
        except subprocess.TimeoutExpired:
            log(f"rocprofv3 --pmc {counter} timed out")
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
    legs = {"host_orient_upload": [], "upload_undirected": [], "device_orient": []}
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        d = g.orient()
        dg = ctx.upload(d)
        t1 = time.perf_counter()
        du = ctx.upload(g)
        t2 = time.perf_counter()
        dd = ctx.orient(du)
        t3 = time.perf_counter()
        if i == 0:
            if dd.download().hash() != d.hash() or not np.array_equal(dd.task_order(), dg.task_order()):
                raise SystemExit("device orientation differs from the host orientation")
        if i >= args.warmup:
            legs["host_orient_upload"].append((t1 - t0) * 1e3)
            legs["upload_undirected"].append((t2 - t1) * 1e3)
            legs["device_orient"].append((t3 - t2) * 1e3)
        log(f"step {i}: host {1e3 * (t1 - t0):.1f} ms, upload {1e3 * (t2 - t1):.1f} ms, "
            f"device orient {1e3 * (t3 - t2):.1f} ms")
        for x in (dd, du, dg):
            x.close()
        del d
    med = {k: round(float(np.median(v)), 3) for k, v in legs.items()}
    dev_total = med["upload_undirected"] + med["device_orient"]
    print(json.dumps({
        "metric": f"host CSR to device DAG on {args.workload}, ms (orientation, reverse lists, task order)",
        "value": round(dev_total, 3), "unit": "ms", "higher_is_better": False, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "config": {"workload": desc, "n": g.n, "undirected_edges": m},
        "legs_ms": med, "host_path_ms": med["host_orient_upload"],
        "speedup_vs_host_path": round(med["host_orient_upload"] / dev_total, 3),
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


def cpu_baseline(g, golden_count, threads):
    import oracle
    if not oracle.ref_available("triangle"):
        return None
    log(f"cpu baseline: reference GraphMiner tc_omp_base, {threads} threads")
    c, t = oracle.ref_count_graph("triangle", g, threads=threads, timeout=600)
    return dict(value=g.nnz // 2 / t, unit="|E|/s", cores=threads, kind="reference",
                sample="full R-MAT-22 seed 2 triangle count (counting loop only)",
                seconds=t, count=c, count_matches_gpu=(c == golden_count))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pattern", default="triangle", choices=list(PATTERN_IDS) + ["motif3", "motif4"])
    ap.add_argument("--workload", default="rmat22", choices=list(WORKLOADS))
    ap.add_argument("--strategy", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic runs")
    ap.add_argument("--sched", default="", choices=["", "compare"],
                    help="clique patterns: time the learned schedule against the heuristic and rank-LUT ones")
    ap.add_argument("--baseline-steps", type=int, default=1)
    ap.add_argument("--prep", default="", choices=["", "compare"],
                    help="time host vs on-device orientation of the workload graph into a device DAG")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="only run the plain copy kernel over 1 GiB (counter calibration)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from conftest import load_xgm
    xgm = load_xgm()
    if args.copy_calibrate:
        ctx = xgm.Context(0)
        print(json.dumps({"copy_bytes": COPY_BYTES, "gbps": ctx.copy_bandwidth(COPY_BYTES, 1)}))
        ctx.close()
        return
    if args.workload == "batch256":
        return bench_batch(args, xgm)
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
        if dom not in per_tier:  # a strategy with one kernel for every edge (bucket bitmap)
            per_tier = {dom: btot}
        ach = per_tier[dom] / (avg[dom] * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None, kernel=dom,
                    alg_bytes_per_launch=per_tier[dom], avg_launch_ms=round(avg[dom], 4))
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
