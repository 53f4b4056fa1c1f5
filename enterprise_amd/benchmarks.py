"""Benchmark modes other than the default plain BFS line of bench.py."""
from __future__ import annotations

import ctypes as C
import time

from ._lib import check, lib


def calib(args):
    """Copy kernel of known size, for scaling hardware byte counters."""
    nbytes = 1 << 30
    ms = C.c_double()
    check(lib().ent_calib_copy(0, nbytes, 5, C.byref(ms)))
    return {"mode": "calib", "bytes_read": nbytes, "bytes_written": nbytes,
            "ms_per_copy": round(ms.value, 4),
            "GBps_read_plus_write": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}


MODES = {"calib": calib}


def run(args):
    t0 = time.perf_counter()
    out = MODES[args.mode](args)
    out.setdefault("wall_s", round(time.perf_counter() - t0, 2))
    return out


def _hmean(xs):
    xs = [x for x in xs if x > 0]
    return len(xs) / sum(1.0 / x for x in xs) if xs else 0.0


def batch(args):
    """Config 4: 512 roots on R-MAT scale 23 (half from the top 1% degree
    vertices, half of degree 1..4) in one multi-source search, against the
    plain single-source path on the same roots. Every batch row is compared
    with the plain result by a device-side row checksum."""
    import sys

    import numpy as np

    from . import _lib as L
    from . import traffic
    from .bfs import DeviceGraph
    from .graph import HostGraph

    def log(m):
        print(m, file=sys.stderr, flush=True)

    scale = getattr(args, "scale", 23) if getattr(args, "scale", 22) != 22 else 23
    k = 512
    t0 = time.perf_counter()
    g = HostGraph.rmat(scale, 16, seed=scale)
    roots = g.pick_roots_mixed(k, k // 2, 1004)
    log(f"[batch] R-MAT scale {scale}: n={g.n} m={g.m} ({time.perf_counter() - t0:.1f}s)")
    dg = DeviceGraph(g, keep_order=getattr(args, "keep_order", False))
    from . import autotune
    tuned = autotune.load(f"rmat{scale}") or autotune.load("rmat22")
    if tuned:
        dg.configure(**{kk: v for kk, v in tuned.items() if kk not in ("reserved", "launch", "hub_cache")})
    # Plain path on the same roots (host window per root, as bench.py plain).
    plain_rates, plain_edges, plain_time, plain_hash = [], 0, 0.0, []
    for r in roots[:8]:
        dg.bfs_device(int(r))
    for r in roots:
        t1 = time.perf_counter()
        st = dg.bfs_device(int(r))
        dt = time.perf_counter() - t1
        plain_edges += st["reached_edges"]
        plain_time += dt
        if st["levels"] > 3:
            plain_rates.append(st["reached_edges"] / dt / 1e9)
        plain_hash.append(int(dg.row_hash(dg.depth_device_ptr(), 1, g.n)[0]))
    log(f"[batch] plain: {len(roots)} roots in {plain_time * 1e3:.1f} ms")
    tuned_batch = autotune.load(f"rmat{scale}_batch")
    if tuned_batch:  # knobs tuned for the multi-source search (the plain runs above use the plain ones)
        dg.configure(**{kk: v for kk, v in tuned_batch.items() if kk in autotune.BATCH_SPACE})
    if getattr(args, "launch", None):  # launch form of the multi-source search only
        dg.configure(launch={"persistent": L.LAUNCH_PERSISTENT, "levels": L.LAUNCH_LEVELS,
                             "graph": L.LAUNCH_GRAPH, "stream": L.LAUNCH_STREAM}[args.launch])
    out = dg.device_alloc(k * g.n * 4)
    try:
        dg.batch_reserve(k, 48)
        if args.warmup > 0:
            dg.bfs_batch_device(roots, out)             # warm-up
        stats = []
        for _ in range(max(1, args.steps if args.steps < 64 else 3)):
            stats.append(dg.bfs_batch_device(roots, out))
        h = dg.row_hash(out, k, g.n)
        mism = int(sum(int(h[i]) != plain_hash[i] for i in range(k)))
    finally:
        dg.device_free(out)
    best = min(stats, key=lambda s: s["host_ms"])
    batch_gteps = best["reached_edges"] / (best["host_ms"] * 1e-3) / 1e9
    plain_agg = plain_edges / plain_time / 1e9
    achieved = best["alg_bytes"] / (best["device_ms"] * 1e-3) / 1e9
    dg.free()
    tr = traffic.figure(f"batch_rmat{scale}")
    return {
        "mode": "batch", "metric": "aggregate GTEPS, 512-source batched vs plain",
        "value": round(batch_gteps, 2), "unit": "GTEPS", "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"multi-source BFS, R-MAT scale {scale} ef 16 (seed {scale}), 512 roots "
                               "(256 top-1% degree, 256 degree 1..4, seed 1004)", "k": k, "words": best["words"],
                   "launch": getattr(args, "launch", None) or "stream"},
        "batch_ms": round(best["host_ms"], 3), "batch_device_ms": round(best["device_ms"], 3),
        "plain_aggregate_gteps": round(plain_agg, 2), "plain_hmean_gteps": round(_hmean(plain_rates), 2),
        "plain_total_ms": round(plain_time * 1e3, 2),
        "speedup_vs_plain": round(batch_gteps / plain_agg, 2),
        "rows_mismatching_plain": mism, "levels": best["levels"], "phases": best["phases"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "alg_bytes": best["alg_bytes"],
                     "traffic": tr, "traffic_over_alg": round(tr / best["alg_bytes"], 3) if tr else None},
        "runs_ms": [round(s["host_ms"], 3) for s in stats],
        "flushes": best["flushes"],
        # The graph and roots, for the CPU baseline that bench.py runs.
        "_cpu_batch": {"graph": g, "roots": roots},
    }


MODES["batch"] = batch


def grid(args):
    """Config 3: 4096 x 4096 4-neighbour grid, 2% of edges deleted (seed
    4096), largest component kept, 16 seeded roots. Compares the launch forms
    on a search of ~8,000 levels: per-phase launches queued on a stream,
    per-phase launches replayed from a captured hipGraph, one persistent
    cooperative launch with grid barriers, and the host-synchronised loop of
    the reference. Every root passes the device all-vertex check; one root per
    launch form is checked against the CPU oracle, which is also the CPU
    baseline (both in bench.py)."""
    import sys

    import numpy as np

    from . import _lib as L
    from .bfs import DeviceGraph
    from .graph import HostGraph

    def log(m):
        print(m, file=sys.stderr, flush=True)

    side = int(getattr(args, "grid_side", 4096) or 4096)
    t0 = time.perf_counter()
    g = HostGraph.grid(side, side, delete_ppm=20000, seed=4096, keep_lcc=True)
    roots = g.pick_roots(16, 1003)
    log(f"[grid] {side}x{side}: n={g.n} m={g.m} ({time.perf_counter() - t0:.1f}s)")
    dg = DeviceGraph(g)
    # Roots timed per form (the slower forms on fewer roots); --grid-roots
    # scales every count down for a shorter run.
    scale_roots = float(getattr(args, "grid_roots", 1.0) or 1.0)
    forms = {name: (launch, max(1, int(round(nr * scale_roots))))
             for name, (launch, nr) in {"stream": (L.LAUNCH_STREAM, 16), "graph": (L.LAUNCH_GRAPH, 16),
                                         "persistent": (L.LAUNCH_PERSISTENT, 4),
                                         "levels": (L.LAUNCH_LEVELS, 2)}.items()}
    only = [f for f in (getattr(args, "grid_forms", "") or "").split(",") if f]
    if only:
        forms = {k: v for k, v in forms.items() if k in only}
    res = {}
    extra = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in getattr(args, "set", []) or []}
    for name, (launch, nroots) in forms.items():
        dg.configure(launch=launch, graph_levels=96, **extra)
        dg.bfs_device(int(roots[0]))   # warm-up (graph capture, launch sizing)
        times, levels, phases = [], [], []
        for r in roots[:nroots]:
            t1 = time.perf_counter()
            st = dg.bfs_device(int(r))
            times.append(time.perf_counter() - t1)
            levels.append(st["levels"])
            phases.append(st["phases"])
        ms = float(np.mean(times)) * 1e3
        res[name] = {"roots": nroots, "ms_per_bfs": round(ms, 3),
                     "us_per_level": round(ms * 1e3 / float(np.mean(levels)), 3),
                     "phases_per_level": round(float(np.mean(phases)) / float(np.mean(levels)), 3),
                     "levels_mean": round(float(np.mean(levels)), 1)}
        log(f"[grid] {name}: {res[name]}")
    # Depths of one root per form (a different root each) and the device
    # all-vertex check of every root; the CPU baseline and the parity check
    # against the oracle run in bench.py on the depths handed back under
    # "_cpu_check".
    checks = []
    for j, (name, (launch, _)) in enumerate(forms.items()):
        dg.configure(launch=launch, graph_levels=96, **extra)
        checks.append((int(roots[j]), dg.bfs(int(roots[j]))[0], name))
    dg.configure(launch=L.LAUNCH_STREAM)
    violations = 0
    single_hash, single_levels = [], []
    for r in roots:
        st = dg.bfs_device(int(r))
        violations += sum(dg.validate_device(int(r)))
        single_hash.append(int(dg.row_hash(dg.depth_device_ptr(), 1, g.n)[0]))
        single_levels.append(st["levels"])
    best = min(res, key=lambda k: res[k]["ms_per_bfs"])
    # The same graph as one 64-source batch keeping 250 levels (the most the
    # level ring holds: it flushes to the rows about every 251 levels),
    # against 64 searches at the best form's mean time; four rows checked
    # against the single-source path. Then the 16 roots above as one 16-source
    # batch: every row against its single-source row, and the batch time per
    # root and per level of a root (a multi-source form of the same 16
    # searches, beside the single-source forms' latency per level).
    batch = batch16 = None
    if not getattr(args, "no_grid_batch", False):
        k = 64
        broots = g.pick_roots(k, 1007)
        out = dg.device_alloc(k * g.n * 4)
        try:
            dg.batch_reserve(k, 250)
            dg.bfs_batch_device(broots[:16], out)   # warm-up (first use of the level ring)
            t1 = time.perf_counter()
            st = dg.bfs_batch_device(broots, out)
            bms = (time.perf_counter() - t1) * 1e3
            h = dg.row_hash(out, k, g.n)
            mism = 0
            for i in (0, 21, 42, 63):
                dg.bfs_device(int(broots[i]))
                mism += int(int(dg.row_hash(dg.depth_device_ptr(), 1, g.n)[0]) != int(h[i]))
            k16 = len(roots)
            t1 = time.perf_counter()
            st16 = dg.bfs_batch_device(roots, out)
            b16 = (time.perf_counter() - t1) * 1e3
            h16 = dg.row_hash(out, k16, g.n)
            per_root = b16 / k16
            batch16 = {"k": k16, "roots": "the 16 roots timed above (seed 1003)", "batch_ms": round(b16, 1),
                       "ms_per_root": round(per_root, 3),
                       "us_per_level_per_root": round(per_root * 1e3 / float(np.mean(single_levels)), 3),
                       # the best form's cost per level times the 16 roots' mean depth
                       "speedup_vs_best_form": round(res[best]["us_per_level"] * float(np.mean(single_levels))
                                                     / (per_root * 1e3), 2),
                       "levels": st16["levels"], "flushes": st16["flushes"],
                       "rows_checked": k16,
                       "rows_mismatching_single_source": int(sum(int(h16[i]) != single_hash[i]
                                                                 for i in range(k16)))}
            log(f"[grid] the 16 roots as one batch: {batch16}")
        finally:
            dg.device_free(out)
            dg.batch_release()
        batch = {"k": k, "roots_seed": 1007, "batch_ms": round(bms, 1), "levels": st["levels"],
                 "flushes": st["flushes"], "kept_levels": 250,
                 "speedup_vs_plain": round(k * res[best]["ms_per_bfs"] / bms, 2),
                 "rows_checked": 4, "rows_mismatching_plain": mism}
        log(f"[grid] batch of {k}: {batch}")
    dg.free()
    return {"mode": "grid", "metric": "us per BFS level, high-diameter grid", "unit": "us/level",
            "value": res[best]["us_per_level"], "best_form": best, "higher_is_better": False,
            "config": {"workload": f"{side}x{side} 4-neighbour grid, 2% edges deleted (seed 4096), "
                                   "largest component, 16 roots (seed 1003)", "n": g.n, "m": g.m},
            "forms": res, "device_validation": {"roots": len(roots), "violations": int(violations)},
            "batch": batch, "batch16": batch16,
            "_cpu_check": {"graph": g, "checks": checks}}


MODES["grid"] = grid


def serve(args):
    """Config 5: a continuous-batching server over two resident graphs, R-MAT
    scale 26 (80% of queries) and scale 20, queries arriving as a seeded
    Poisson stream at each rate, uniform non-isolated roots. Each graph's
    first slot takes batches of up to --serve-batch queries (default 512),
    the others up to --serve-small-batch (default 64) while the queue fits
    in them; every slot keeps --serve-levels levels between flushes
    (default 8; R-MAT 26 searches rarely need more). Reports p50/p99 latency
    (submit to end of the serving batch) and GTEPS per rate; sampled queries
    are checked against the single-source path by device row checksums."""
    import sys

    import numpy as np

    from .bfs import DeviceGraph
    from .graph import HostGraph
    from .serving import BFSServer, run_load

    def log(m):
        print(m, file=sys.stderr, flush=True)

    big_scale = int(getattr(args, "serve_scale", 26) or 26)
    rates = [int(x) for x in str(getattr(args, "rates", "100,200,500,1000,2000,5000")).split(",")]
    dur = float(getattr(args, "duration", 2.0) or 2.0)
    t0 = time.perf_counter()
    gb = HostGraph.rmat(big_scale, 16, seed=big_scale)
    log(f"[serve] R-MAT {big_scale}: n={gb.n} m={gb.m} ({time.perf_counter() - t0:.1f}s)")
    gs = HostGraph.rmat(20, 16, seed=20)
    roots = {"large": gb.pick_roots(4096, 1005), "small": gs.pick_roots(4096, 1006)}
    db, ds = DeviceGraph(gb), DeviceGraph(gs)
    from . import autotune
    tuned = autotune.load("rmat22")
    if tuned:
        for d in (db, ds):
            d.configure(**{k: v for k, v in tuned.items() if k not in ("reserved", "launch", "hub_cache")})
    m_big = gb.m
    del gb  # the host copy is no longer needed
    small = int(getattr(args, "serve_small_batch", 64) or 0)
    srv = BFSServer(mem_budget_bytes=int(getattr(args, "serve_budget_gb", 256)) << 30,
                    max_batch=int(getattr(args, "serve_batch", 512)), max_levels=int(getattr(args, "serve_levels", 8)),
                    max_slots=int(getattr(args, "serve_slots", 3)), plain_max=int(getattr(args, "plain_max", 12)),
                    small_batch=small or None)
    slots_big = srv.add_graph("large", db, share_budget=0.9)
    slots_small = srv.add_graph("small", ds, share_budget=0.1)
    log(f"[serve] slots: large {slots_big}, small {slots_small}")
    srv.start()
    run_load(srv, 200, 0.3, 1, roots, {"large": 0.8, "small": 0.2})   # warm-up
    per_rate = []
    try:
        for rate in rates:
            for lane in srv.lanes.values():
                lane.stats.update({"batches": 0, "plain_batches": 0, "queries": 0, "reached_edges": 0,
                                   "alg_bytes": 0, "busy_s": 0.0, "batch_sizes": [], "flushes": 0})
            t1 = time.perf_counter()
            res = run_load(srv, rate, dur, 1005 + rate, roots, {"large": 0.8, "small": 0.2})
            wall = time.perf_counter() - t1
            lat = np.array([q.latency_ms for q in res])
            edges = sum(l.stats["reached_edges"] for l in srv.lanes.values())
            alg = sum(l.stats["alg_bytes"] for l in srv.lanes.values())
            sizes = sum((l.stats["batch_sizes"] for l in srv.lanes.values()), [])
            # Sampled check against the single-source path.
            rng = np.random.default_rng(rate)
            mism = 0
            sample = rng.choice(len(res), size=min(8, len(res)), replace=False)
            for i in sample:
                q = res[int(i)]
                d = db if q.graph == "large" else ds
                d.bfs_device(q.root)
                mism += int(int(d.row_hash(d.depth_device_ptr(), 1, d.n)[0]) != q.row_hash)
            row = {"rate_qps": rate, "queries": len(res), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                   "p99_ms": round(float(np.percentile(lat, 99)), 2),
                   "throughput_qps": round(len(res) / wall, 1), "gteps": round(edges / wall / 1e9, 2),
                   "alg_GBps": round(alg / wall / 1e9, 1),
                   "mean_batch": round(float(np.mean(sizes)), 1) if sizes else 0,
                   "flushes": sum(l.stats["flushes"] for l in srv.lanes.values()),
                   "checked": int(len(sample)), "mismatches": mism}
            log(f"[serve] {row}")
            per_rate.append(row)
    finally:
        srv.close()
        db.free()
        ds.free()
    return {"mode": "serve", "metric": "p50/p99 query latency vs GTEPS", "unit": "ms",
            "config": {"workload": f"serving R-MAT {big_scale} (80%) + R-MAT 20, Poisson arrivals, "
                                   f"continuous batching of up to {srv.max_batch} queries", "m_large": int(m_big),
                       "max_batch": srv.max_batch, "small_batch": srv.small_batch, "kept_levels": srv.max_levels,
                       "slots": {"large": slots_big, "small": slots_small}, "duration_s": dur,
                       "plain_max": srv.plain_max},
            "rates": per_rate}


MODES["serve"] = serve


def csr(args):
    """CSR construction on the device (include/ent_convert.h) from the raw
    tuples of the R-MAT scale-S graph (ef 16, seed S, relabelled): both
    directions, loops dropped, rows sorted, duplicates dropped
    (ENT_CSR_SIMPLE), with the tuples resident in HBM when timed. The result
    is compared bit for bit with the host generator's CSR. CPU figures: the
    library's multi-threaded host builder on the same tuples, and (in
    bench.py) the oracle's single-core construction on a prefix of them."""
    import ctypes as C
    import sys

    import numpy as np

    import os

    from . import convert
    from .graph import HostGraph

    def log(m):
        print(m, file=sys.stderr, flush=True)

    scale = int(getattr(args, "scale", 22) or 22)
    n = 1 << scale
    t0 = time.perf_counter()
    src, dst = convert.rmat_tuples(scale, 16, seed=scale)
    ne = int(src.size)
    log(f"[csr] R-MAT scale {scale}: {ne} tuples ({time.perf_counter() - t0:.1f}s)")
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    bufs = []

    def dmalloc(nbytes):
        p = C.c_void_p()
        if hip.hipMalloc(C.byref(p), nbytes) != 0:
            raise RuntimeError(f"hipMalloc({nbytes}) failed")
        bufs.append(p.value)
        return p.value

    try:
        ds, dd = dmalloc(4 * ne), dmalloc(4 * ne)
        db, da = dmalloc(8 * (n + 1)), dmalloc(4 * 2 * ne)
        assert hip.hipMemcpy(ds, src.ctypes.data, 4 * ne, 1) == 0
        assert hip.hipMemcpy(dd, dst.ctypes.data, 4 * ne, 1) == 0
        ms, st = [], None
        for k in range(max(1, args.warmup) + max(1, args.steps)):
            m, st = convert.csr_build_device(n, ne, ds, dd, convert.SIMPLE, db, da, 2 * ne)
            if k >= max(1, args.warmup):
                ms.append(st["device_ms"])
        beg = np.empty(n + 1, dtype=np.int64)
        adj = np.empty(m, dtype=np.int32)
        assert hip.hipMemcpy(beg.ctypes.data, db, 8 * (n + 1), 2) == 0
        assert hip.hipMemcpy(adj.ctypes.data, da, 4 * m, 2) == 0
    finally:
        for p in bufs:
            hip.hipFree(p)
    t1 = time.perf_counter()
    h = HostGraph.from_edges(n, src, dst)
    host_s = time.perf_counter() - t1
    parity = bool(np.array_equal(beg, h.beg_pos) and np.array_equal(adj, h.adj))
    dev_ms = float(np.median(ms))
    alg = 8 * ne + 4 * m + 8 * (n + 1)
    passes = (st["key_bits"] + 7) // 8
    return {"mode": "csr", "metric": "CSR construction from raw tuples, million tuples/s", "unit": "Mtuples/s",
            "value": round(ne / (dev_ms * 1e3), 1), "higher_is_better": True, "ms": round(dev_ms, 3),
            "ms_all": [round(x, 3) for x in ms],
            "config": {"workload": f"R-MAT scale {scale} ef 16 raw tuples (seed {scale}, relabelled) -> "
                                   "symmetric CSR, loops and duplicates dropped, rows sorted",
                       "n": n, "tuples": ne, "m": int(m), "records": int(st["records"]),
                       "key_bits": int(st["key_bits"]), "radix_passes": int(passes)},
            "parity_vs_host_generator": parity,
            "roofline": {"bound": "hbm", "achieved": round(alg / (dev_ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                         "unit": "GB/s", "frac": round(alg / (dev_ms * 1e-3) / 8e12, 4),
                         "traffic": None,
                         "algorithmic_bytes": int(alg),
                         "sort_bytes_estimate": int(st["records"]) * 2 * 8 * passes},
            "cpu_host_builder": {"value": round(ne / (host_s * 1e6), 1), "unit": "Mtuples/s",
                                 "threads": min(os.cpu_count() or 1, 16),
                                 "sample": f"library host builder (ent_csr_from_edges) on the same {ne} tuples",
                                 "ms": round(host_s * 1e3, 1)},
            "_cpu_csr": {"n": n, "src": src, "dst": dst}}


MODES["csr"] = csr
