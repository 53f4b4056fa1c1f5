"""Benchmark of the MI355X direction-optimizing BFS.

Default workload: R-MAT scale 22, edgefactor 16 (A/B/C/D = .57/.19/.19/.05,
seed 22, relabelled, symmetrized, loops and duplicates removed), 64 seeded
non-isolated roots (seed 1002) run one at a time on the resident graph. One
step is one sweep over all 64 roots; --steps K times K sweeps. Each search is
timed on the host around the C-ABI call (the reference's window,
bfs_gpu_opt.cuh:381-411: from the call to the end of the search, depth array
left in HBM). A root's time is its mean over the K sweeps; the reported value
is the harmonic mean over the roots of per-root GTEPS with the reference's
numerator M_c = sum of degrees over reached vertices (validate.h:115-123),
counting, as the reference does, the roots whose search goes deeper than two
levels (bfs_gpu_opt.cuh:413-445).

The default line also carries, under "extra", the 512-source multi-source
search on R-MAT scale 23 against the plain path on the same roots
(extra.batch) and the per-level cost of each launch form on the 4096^2
high-diameter grid, with the same grid as one 64-source batch and its 16
roots as one 16-source batch (extra.grid), each with its oracle parity and
CPU figures (--no-extra skips both).

Other modes (--mode): grid, batch, serve (continuous-batching serving loop)
and csr (CSR construction on the device from the raw tuples of R-MAT scale
--scale); each prints its own JSON line.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "GTEPS (hmean, 64 roots) plain vs batched; HBM GB/s; p99 query ms, 1 MI355X"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def hmean(xs):
    xs = [x for x in xs if x > 0]
    return len(xs) / sum(1.0 / x for x in xs) if xs else 0.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline_serial(g, roots, budget_s, gpu_depths=None):
    """Serial queue BFS of the oracle on one host core over a bounded sample
    of the same roots; also checks the GPU depths of those roots."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ct
    oracle = oracle_ct.Oracle()
    rates, checked, mismatches = [], 0, 0
    t_start = time.perf_counter()
    for i, r in enumerate(roots):
        t0 = time.perf_counter()
        d = oracle.bfs(g.beg_pos, g.adj, int(r))
        dt = time.perf_counter() - t0
        edges, _ = oracle.report(g.beg_pos, d)
        if d.max() > 2:
            rates.append(edges / dt / 1e9)
        if gpu_depths is not None and i in gpu_depths:
            checked += 1
            mismatches += int(not np.array_equal(d, gpu_depths[i]))
        if time.perf_counter() - t_start > budget_s:
            break
    return {"value": round(hmean(rates), 4), "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": f"serial queue BFS (oracle) on {len(rates)} of the {len(roots)} roots, "
                      f"R-MAT scale {g.scale_tag}, host {cpu_model()}, "
                      f"{os.cpu_count()} host cores visible"}, checked, mismatches


def grid_cpu_check(g, checks):
    """Grid mode: the serial oracle's time per BFS (the CPU baseline) and
    parity of the GPU depths it was handed, on the same roots."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ct
    oracle = oracle_ct.Oracle()
    cpu_ms, mism, forms = [], 0, []
    for r, got, form in checks:
        t0 = time.perf_counter()
        want = oracle.bfs(g.beg_pos, g.adj, int(r))
        cpu_ms.append((time.perf_counter() - t0) * 1e3)
        mism += int(not np.array_equal(got, want))
        forms.append(form)
    return {"cpu_serial_ms_per_bfs": round(float(np.mean(cpu_ms)), 1), "parity_roots": len(checks),
            "parity_forms": forms, "parity_mismatches": mism}


def csr_cpu_baseline(n, src, dst, sample=64 << 20):
    """CSR mode: the oracle's single-core construction (same flags: both
    directions, no loops, sorted deduplicated rows) on the first `sample`
    tuples, in million tuples per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ct
    oracle = oracle_ct.Oracle()
    k = min(sample, int(src.size))
    t0 = time.perf_counter()
    oracle.csr_build(n, src[:k], dst[:k], 1 | 2 | 8)
    dt = time.perf_counter() - t0
    return {"value": round(k / dt / 1e6, 2), "unit": "Mtuples/s", "cores": 1, "kind": "port",
            "sample": f"oracle construction (restated converter, sorted and deduplicated) on the first {k} "
                      f"of {src.size} tuples", "ms": round(dt * 1e3, 1)}


def cpu_baseline_openmp(g, roots, budget_s, threads):
    """Direction-optimizing OpenMP BFS of the oracle on `threads` host cores
    over a bounded sample of the same roots."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ct
    oracle = oracle_ct.Oracle()
    rates = []
    t_start = time.perf_counter()
    for r in roots:
        t0 = time.perf_counter()
        d = oracle.bfs_do(g.beg_pos, g.adj, int(r), threads=threads)
        dt = time.perf_counter() - t0
        edges, _ = oracle.report(g.beg_pos, d)
        if d.max() > 2:
            rates.append(edges / dt / 1e9)
        if time.perf_counter() - t_start > budget_s:
            break
    return {"value": round(hmean(rates), 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"OpenMP direction-optimizing BFS (oracle) on {threads} threads, {len(rates)} of the "
                      f"{len(roots)} roots, R-MAT scale {g.scale_tag}, host {cpu_model()}"}


def host_threads():
    try:
        return max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "0")) or 1 << 30))
    except (AttributeError, ValueError):
        return os.cpu_count() or 1


def load_tuned(dg, name):
    path = os.path.join(ROOT, "enterprise_amd", "tuned", f"{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            cfg = json.load(f)
        dg.configure(**{k: v for k, v in cfg.get("config", cfg).items() if k not in ("reserved", "launch", "hub_cache")})
        return path
    return None


def batch_cpu_baseline(g, roots, budget_s, threads):
    """Batch mode: the OpenMP direction-optimizing oracle BFS on `threads`
    host cores over a bounded prefix of the same 512 roots (one at a time),
    as aggregate GTEPS (sum of M_c over sum of times)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ct
    oracle = oracle_ct.Oracle()
    edges_sum, t_sum, done = 0, 0.0, 0
    for r in roots:
        t0 = time.perf_counter()
        d = oracle.bfs_do(g.beg_pos, g.adj, int(r), threads=threads)
        t_sum += time.perf_counter() - t0
        edges_sum += oracle.report(g.beg_pos, d)[0]
        done += 1
        if t_sum > budget_s:
            break
    return {"value": round(edges_sum / t_sum / 1e9, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"OpenMP direction-optimizing BFS (oracle) on {threads} threads, the first {done} of the "
                      f"{len(roots)} roots one at a time, aggregate; host {cpu_model()}"}


def run_extra(args, mode, **over):
    """Runs benchmarks mode `mode` with argument overrides and its CPU legs
    (oracle parity and baseline, which stay in this file)."""
    from enterprise_amd import benchmarks
    ns = argparse.Namespace(**vars(args))
    for k, v in over.items():
        setattr(ns, k, v)
    ns.mode = mode
    out = benchmarks.run(ns)
    chk = out.pop("_cpu_check", None)
    if chk is not None:
        out.update(grid_cpu_check(chk["graph"], chk["checks"]))
    cb = out.pop("_cpu_batch", None)
    if cb is not None:
        out["cpu_baseline"] = None if args.no_cpu else batch_cpu_baseline(cb["graph"], cb["roots"],
                                                                           args.cpu_budget / 2, host_threads())
    return out


def bench_plain(args):
    from enterprise_amd import DeviceGraph, HostGraph, traffic
    from enterprise_amd import _lib as L

    scale = args.scale
    t0 = time.perf_counter()
    g = HostGraph.rmat(scale, 16, seed=scale)
    g.scale_tag = scale
    log(f"[bench] R-MAT scale {scale}: n={g.n} m={g.m} generated in {time.perf_counter() - t0:.1f}s")
    roots = g.pick_roots(64, 1002)
    t0 = time.perf_counter()
    dg = DeviceGraph(g, device=0, keep_order=args.keep_order)
    upload_s = time.perf_counter() - t0
    tuned = load_tuned(dg, f"rmat{scale}")
    if args.set:
        dg.configure(**{kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.set})
    if args.launch is not None:
        dg.configure(launch={"persistent": L.LAUNCH_PERSISTENT, "levels": L.LAUNCH_LEVELS,
                             "graph": L.LAUNCH_GRAPH, "stream": L.LAUNCH_STREAM}[args.launch])
    log(f"[bench] config {dg.config.as_dict()} (tuned file: {tuned})")
    for _ in range(args.warmup):              # one warm-up step = one sweep over the roots
        for r in roots:
            dg.bfs_device(int(r))
    log("[bench] warmup done")
    host_s = [[] for _ in roots]              # per root: host window of each sweep
    ev_ms = [[] for _ in roots]               # per root: HIP-event device time of each sweep
    stats = [None] * len(roots)
    t_all0 = time.perf_counter()
    for _ in range(args.steps):
        for i, r in enumerate(roots):
            t0 = time.perf_counter()
            st = dg.bfs_device(int(r))
            host_s[i].append(time.perf_counter() - t0)
            # HIP-event device time of the same search, read after the host
            # window (it waits for the launches queued behind the finished
            # search).
            ev_ms[i].append(dg.last_device_ms())
            stats[i] = st
    t_all = time.perf_counter() - t_all0
    log(f"[bench] timed {args.steps} sweeps of {len(roots)} roots in {t_all:.3f}s")
    # Depths of a few timed roots for the CPU check (outside the timed region);
    # these calls copy the depth array to host memory, so their wall time is
    # the PCIe-inclusive rate.
    pcie = []
    gpu_depths = {}
    for i in range(min(args.check_roots, 64)):
        t0 = time.perf_counter()
        d, st = dg.bfs(int(roots[i]))
        pcie.append(st["reached_edges"] / (time.perf_counter() - t0) / 1e9)
        gpu_depths[i] = d
    # Every timed root once more, each result checked on the device for all
    # vertices (ent_bfs_validate_device; outside the timed region).
    violations = 0
    for r in roots:
        dg.bfs_device(int(r))
        violations += sum(dg.validate_device(int(r)))
    # Schedule-exact examined bytes: the counting kernel variant over the
    # same roots (outside the timed region; it is slower).
    exam = []
    if not args.no_examined:
        dg.configure(count_examined=1)
        exam = [dg.bfs_device(int(r))["examined_bytes"] for r in roots]
        dg.configure(count_examined=0)
    used = [i for i in range(len(roots)) if stats[i]["levels"] - 1 > 2]
    mean_s = {i: float(np.mean(host_s[i])) for i in used}
    gteps = [stats[i]["reached_edges"] / mean_s[i] / 1e9 for i in used]
    gteps_dev = [stats[i]["reached_edges"] / (float(np.mean(ev_ms[i])) * 1e-3) / 1e9 for i in used]
    alg_bytes = [stats[i]["alg_bytes"] for i in used]
    # Algorithmic bytes over HIP-event device time, every timed search of the
    # counted roots.
    dev_sum_s = sum(sum(ev_ms[i]) for i in used) * 1e-3
    achieved = sum(a * args.steps for a in alg_bytes) / dev_sum_s / 1e9
    host_sum_s = sum(sum(host_s[i]) for i in used)
    achieved_host = sum(a * args.steps for a in alg_bytes) / host_sum_s / 1e9
    value = hmean(gteps)
    all_host_ms = [x * 1e3 for i in used for x in host_s[i]]
    cpu = None
    cpu_serial = None
    parity = None
    if not args.no_cpu:
        cpu_serial, checked, mism = cpu_baseline_serial(g, roots, args.cpu_budget / 2, gpu_depths)
        parity = {"roots_checked": checked, "mismatches": mism}
        cpu = cpu_baseline_openmp(g, roots, args.cpu_budget / 2, host_threads())
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_all * 1e3 / max(1, args.steps), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"plain single-source DO-BFS, R-MAT scale {scale} ef 16 "
                               f"(a=.57 b=.19 c=.19, seed {scale}), 64 roots (seed 1002); "
                               "one step = one sweep over the 64 roots",
                   "n": g.n, "m_directed": g.m, "roots": len(roots), "roots_in_mean": len(used),
                   "sweeps": args.steps,
                   "launch": dg.config.launch, "direction": dg.config.direction,
                   "rows": "caller order" if args.keep_order else "by neighbour degree class"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic.figure(f"plain_rmat{scale}"),
                     "kernel": "whole search (init + level kernels) per BFS, HIP events on "
                               "the handle's stream",
                     "alg_bytes_per_bfs_mean": int(np.mean(alg_bytes)),
                     "examined_bytes_per_bfs_mean": int(np.mean(exam)) if exam else None,
                     "examined_over_alg": round(float(np.mean(exam)) / float(np.mean(alg_bytes)), 3) if exam else None},
        "cpu_baseline": cpu,
        "extra": {
            "gteps_hmean_device_time": round(hmean(gteps_dev), 3),
            "gteps_amean": round(float(np.mean(gteps)), 3),
            "gteps_graph500_undirected_hmean": round(value / 2, 3),
            "alg_GBps_host_window": round(achieved_host, 1),
            "ms_per_bfs_host_window_mean": round(float(np.mean(all_host_ms)), 4),
            "device_ms_mean": round(float(np.mean([x for i in used for x in ev_ms[i]])), 4),
            "device_clock_ms_mean": round(float(np.mean([stats[i]["device_ms"] for i in used])), 4),
            "host_ms_p50": round(float(np.percentile(all_host_ms, 50)), 4),
            "levels_mean": round(float(np.mean([stats[i]["levels"] for i in used])), 2),
            "upload_s": round(upload_s, 3),
            "parity": parity,
            "device_validation": {"roots": len(roots), "violations": int(violations)},
            "cpu_serial_1core": cpu_serial,
            "gteps_hmean_pcie_inclusive": round(hmean(pcie), 3) if pcie else None,
        },
    }
    dg.free()
    del g
    if not args.no_extra:
        log("[bench] extra: 512-source batch on R-MAT scale 23")
        b = run_extra(args, "batch", scale=23, steps=3, warmup=1)
        out["extra"]["batch"] = {k: b[k] for k in (
            "value", "unit", "batch_ms", "batch_device_ms", "plain_aggregate_gteps", "speedup_vs_plain",
            "rows_mismatching_plain", "levels", "flushes", "roofline", "cpu_baseline", "config")}
        log(f"[bench] batch: {out['extra']['batch']['speedup_vs_plain']}x the plain path")
        log("[bench] extra: 4096^2 grid, launch forms")
        gr = run_extra(args, "grid", grid_roots=0.25, set=[])
        out["extra"]["grid"] = {k: gr[k] for k in (
            "value", "unit", "best_form", "forms", "device_validation", "batch", "batch16", "cpu_serial_ms_per_bfs",
            "parity_roots", "parity_forms", "parity_mismatches", "config") if k in gr}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="plain mode: sweeps over the 64 roots (one step = one sweep)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-grid-batch", action="store_true",
                    help="grid mode: skip the 64-source batch of the grid")
    ap.add_argument("--no-extra", action="store_true",
                    help="plain mode: skip the batch and grid figures after the plain line")
    ap.add_argument("--grid-roots", type=float, default=1.0,
                    help="grid mode: fraction of the roots timed per launch form")
    ap.add_argument("--mode", default="plain", choices=["plain", "grid", "batch", "serve", "calib", "csr"])
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--launch", default=None, choices=["persistent", "levels", "graph", "stream"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--check-roots", type=int, default=4)
    ap.add_argument("--no-examined", action="store_true",
                    help="skip the examined-work pass (counting kernel variant) after the timed searches")
    ap.add_argument("--serve-scale", type=int, default=26)
    ap.add_argument("--serve-budget-gb", type=int, default=256,
                    help="serve mode: device memory for the slots, GiB (the MI355X has 288 GB)")
    ap.add_argument("--serve-batch", type=int, default=512,
                    help="serve mode: queries per batch on each graph's first slot")
    ap.add_argument("--serve-small-batch", type=int, default=64,
                    help="serve mode: queries per batch on the other slots (0: all slots --serve-batch)")
    ap.add_argument("--serve-levels", type=int, default=8,
                    help="serve mode: levels a slot keeps between flushes to the rows")
    ap.add_argument("--serve-slots", type=int, default=3, help="serve mode: most slots per graph")
    ap.add_argument("--rates", default="100,200,500,1000,2000,5000")
    ap.add_argument("--duration", type=float, default=2.0)
    ap.add_argument("--plain-max", type=int, default=12,
                    help="serve: batches of at most this many queries run as single-source searches")
    ap.add_argument("--grid-side", type=int, default=4096)
    ap.add_argument("--grid-forms", default="", help="comma-separated launch forms for --mode grid "
                                                       "(stream,graph,persistent,levels; default all)")
    ap.add_argument("--traffic", default=None, metavar="PATH",
                    help="measure HBM bytes per search of --mode (plain or batch) with two "
                         "rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) scaled by a copy kernel; "
                         "writes the summary to PATH")
    ap.add_argument("--keep-order", action="store_true",
                    help="upload the adjacency rows in the caller's order (default: by neighbour degree class)")
    ap.add_argument("--set", action="append", default=[],
                    help="config override key=value (ent_bfs_config field), repeatable")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(batched and serving work would shard by replicating the graph across GPUs)",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic:
        from enterprise_amd import traffic
        if args.mode == "plain":
            child = ["--no-cpu", "--no-examined", "--no-extra", "--steps", "1", "--warmup", "0",
                     "--check-roots", "0", "--scale", str(args.scale)]
        elif args.mode == "batch":
            child = ["--mode", "batch", "--steps", "1", "--warmup", "0", "--no-cpu"]
        else:
            raise SystemExit("--traffic supports --mode plain and batch")
        child += [x for kv in args.set for x in ("--set", kv)]
        out = traffic.measure(args.mode, child, os.path.join(os.path.dirname(os.path.abspath(args.traffic)),
                                                             "traffic_runs"))
        with open(args.traffic, "w") as f:
            json.dump(out, f, indent=1)
    elif args.mode == "plain":
        out = bench_plain(args)
    else:
        from enterprise_amd import benchmarks
        out = run_extra(args, args.mode) if args.mode in ("grid", "batch") else benchmarks.run(args)
        csr = out.pop("_cpu_csr", None)
        if csr is not None:
            out["cpu_baseline"] = csr_cpu_baseline(csr["n"], csr["src"], csr["dst"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
