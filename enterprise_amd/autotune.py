"""Autotuner for the BFS launch and classification parameters.

Coordinate descent over the ent_bfs_config knobs (degree-class thresholds,
block size, blocks per CU, bottom-up probe length and chunk size, the
direction-switch alpha/beta, launch form): each candidate runs a sample of
roots on the resident graph and is scored by the harmonic mean of per-root
GTEPS (host window around the C-ABI call, as the benchmark measures). A
candidate is accepted only if its depth arrays equal those of the starting
configuration on check roots (the depths do not depend on the schedule, so
any difference is a bug, not a tuning result). The winner is written as JSON
and loaded by bench.py / the serving loop for graphs of the same family.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _lib as L
from .bfs import DeviceGraph

TUNED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned")

# Candidate values per knob (every value satisfies the library's checks given
# the ordering constraints, which are enforced before a candidate runs).
SPACE = {
    "block_size": [128, 256, 512],
    "blocks_per_cu": [2, 4, 8, 12],
    "t_wave": [8, 16, 32, 64],
    "t_block": [256, 512, 1024, 2048, 4096],
    "t_grid": [1024, 2048, 4096, 8192, 16384],
    "piece": [512, 1024, 2048, 4096],
    "bu_probe": [4, 8, 16, 32, 64],
    "bu_chunk_words": [2, 4, 8, 16, 32],
    "alpha": [4, 8, 15, 30, 60],
    "beta": [6, 12, 18, 36, 72],
    "queue_small": [-1, 1024, 4096, 16384, 65536],
    "queue_max_degree": [0, 64, 256, 4096, 32768],
    "hub_cache": [0, 8, 10, 12],
    "hub_td": [0, 1],
}


def _valid(cfg):
    return (1 <= cfg["t_wave"] <= cfg["t_block"] <= cfg["t_grid"]
            and 64 <= cfg["piece"] <= cfg["t_grid"])


def _hmean(xs):
    xs = [x for x in xs if x > 0]
    return len(xs) / sum(1.0 / x for x in xs) if xs else 0.0


def measure(dg, roots, reps=3):
    """Harmonic mean over roots of GTEPS, each root's best of `reps` runs
    (host window around the C-ABI call, as the benchmark measures)."""
    best = {}
    for _ in range(reps):
        for r in roots:
            t0 = time.perf_counter()
            st = dg.bfs_device(int(r))
            dt = time.perf_counter() - t0
            if st["levels"] > 3:
                rate = st["reached_edges"] / dt / 1e9
                best[int(r)] = max(best.get(int(r), 0.0), rate)
    return _hmean(list(best.values()))


def tune(dg: DeviceGraph, roots, check_roots, rounds=2, log=print, budget_s=120.0, space=None, reps=3):
    """Returns (best config dict, best score, history). space: knob -> list
    of candidate values (default SPACE)."""
    space = SPACE if space is None else space
    base = dg.config.as_dict()
    want = {int(r): dg.bfs(int(r))[0] for r in check_roots}
    best = dict(base)
    for r in roots[:4]:
        dg.bfs_device(int(r))  # warm up
    best_score = measure(dg, roots, reps)
    history = [(dict(best), best_score)]
    log(f"[tune] start {best_score:.1f} GTEPS")
    t_start = time.perf_counter()
    for rnd in range(rounds):
        improved = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob] or time.perf_counter() - t_start > budget_s:
                    continue
                cand = dict(best)
                cand[knob] = v
                if not _valid(cand):
                    continue
                try:
                    dg.configure(**cand)
                except L.EntError:
                    continue
                ok = all(np.array_equal(dg.bfs(r)[0], d) for r, d in want.items())
                if not ok:
                    raise RuntimeError(f"depth mismatch under config {cand}")
                score = measure(dg, roots, reps)
                history.append((dict(cand), score))
                if score > best_score * 1.01:
                    log(f"[tune] {knob}={v}: {score:.1f} GTEPS (was {best_score:.1f})")
                    best, best_score, improved = cand, score, True
        if not improved:
            break
    dg.configure(**best)
    return best, best_score, history


# Knobs of the multi-source search (config fields it reads).
BATCH_SPACE = {
    "batch_alpha": [1, 2, 4, 8, 16],
    "bu_probe": [4, 8, 16, 32, 64],
    "t_wave": [8, 16, 32, 64],
    "t_block": [256, 512, 1024, 2048],
    "bu_chunk_words": [2, 4, 8, 16],
    "blocks_per_cu": [4, 8, 12, 16],
}


def measure_batch(dg, roots, out_ptr, reps=1):
    """Aggregate GTEPS of one multi-source search over all roots (host window)."""
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        st = dg.bfs_batch_device(roots, out_ptr)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return st["reached_edges"] / best / 1e9


def tune_batch(dg: DeviceGraph, roots, rounds=2, log=print, budget_s=150.0, space=None, reps=3):
    """Coordinate descent over BATCH_SPACE (or `space`) for one multi-source
    search over `roots`, scored by aggregate GTEPS; a candidate must
    reproduce every row of the starting configuration (device row
    checksums)."""
    space = BATCH_SPACE if space is None else space
    k = len(roots)
    out = dg.device_alloc(k * dg.n * 4)
    try:
        dg.batch_reserve(k, 0)
        base = dg.config.as_dict()
        dg.bfs_batch_device(roots, out)  # warm-up
        want = dg.row_hash(out, k, dg.n).copy()
        best = dict(base)
        best_score = measure_batch(dg, roots, out, reps=reps)
        history = [(dict(best), best_score)]
        log(f"[tune-batch] start {best_score:.1f} GTEPS")
        t_start = time.perf_counter()
        for rnd in range(rounds):
            improved = False
            for knob, values in space.items():
                for v in values:
                    if v == best[knob] or time.perf_counter() - t_start > budget_s:
                        continue
                    cand = dict(best)
                    cand[knob] = v
                    if not _valid(cand):
                        continue
                    try:
                        dg.configure(**cand)
                    except L.EntError:
                        continue
                    score = measure_batch(dg, roots, out, reps=reps)
                    if not np.array_equal(dg.row_hash(out, k, dg.n), want):
                        raise RuntimeError(f"row mismatch under config {cand}")
                    history.append((dict(cand), score))
                    if score > best_score * 1.01:
                        log(f"[tune-batch] {knob}={v}: {score:.1f} GTEPS (was {best_score:.1f})")
                        best, best_score, improved = cand, score, True
            if not improved:
                break
        dg.configure(**best)
        return best, best_score, history
    finally:
        dg.device_free(out)


def save(name, cfg, score, meta, out_dir=None):
    d = TUNED_DIR if out_dir is None else out_dir
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"{name}.json")
    with open(path, "w") as f:
        json.dump({"config": cfg, "score_gteps_hmean": round(score, 2), "meta": meta}, f, indent=1)
    return path


def load(name, tuned_dir=None):
    path = os.path.join(TUNED_DIR if tuned_dir is None else tuned_dir, f"{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["config"]


def main():
    import argparse
    import sys
    from .graph import HostGraph
    ap = argparse.ArgumentParser(description="tune the BFS configuration on an R-MAT graph")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--roots", type=int, default=32)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--budget", type=float, default=150.0)
    ap.add_argument("--out", default=None, help="output directory for the JSON")
    ap.add_argument("--seed", type=int, default=None,
                    help="R-MAT seed of the tuning graph (default: scale + 100, a graph other than "
                         "the benchmark's, whose seed is the scale)")
    ap.add_argument("--from-tuned", action="store_true",
                    help="start the descent from the saved configuration for this graph, if any")
    ap.add_argument("--batch", type=int, default=0,
                    help="tune the multi-source search with this many roots (half of them from the "
                         "top-1%% degree vertices) instead of the single-source search")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    seed = args.scale + 100 if args.seed is None else args.seed
    g = HostGraph.rmat(args.scale, 16, seed=seed)
    dg = DeviceGraph(g)
    dg.configure(launch=L.LAUNCH_STREAM)
    if args.batch:
        roots = g.pick_roots_mixed(args.batch, args.batch // 2, 777)
        best, score, hist = tune_batch(dg, roots, rounds=args.rounds, log=log, budget_s=args.budget)
        name = f"rmat{args.scale}_batch"
        meta = {"graph": f"rmat{args.scale} ef16 seed {seed}", "roots": args.batch,
                "roots_kind": "half top-1% degree, half degree 1..4 (seed 777)", "candidates": len(hist),
                "score": "aggregate GTEPS of one multi-source search"}
    else:
        roots = g.pick_roots(args.roots, 777)
        start = load(f"rmat{args.scale}") if args.from_tuned else None
        if start:
            dg.configure(**{k: v for k, v in start.items() if k not in ("reserved", "launch")})
        best, score, hist = tune(dg, roots, roots[:3], rounds=args.rounds, log=log, budget_s=args.budget)
        name = f"rmat{args.scale}"
        meta = {"graph": f"rmat{args.scale} ef16 seed {seed}", "roots": args.roots,
                "candidates": len(hist)}
    path = save(name, best, score, meta, args.out)
    print(json.dumps({"tuned": path, "score": score, "config": best}))


if __name__ == "__main__":
    main()
