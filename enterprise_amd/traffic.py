"""HBM traffic per search from hardware counters.

Runs the benchmark under `rocprofv3 --pmc FETCH_SIZE` and, separately,
`--pmc WRITE_SIZE` (one counter group per run), each as a child process, plus
the same two counters over a copy kernel of known size (ent_calib_copy, 1 GiB
read and 1 GiB written per launch). The counters are reported in KiB per
dispatch; the copy kernel gives the factor from counter units to bytes. The
traffic of one search is the scaled sum over the search's kernels divided by
the number of searches (one k_init / k_binit_zero dispatch per search).
"""
from __future__ import annotations

import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Kernels of one search, by mode: (kernels counted, kernel that starts a search).
SEARCH_KERNELS = {
    "plain": (("ent::k_init", "ent::k_level"), "ent::k_init"),
    "batch": (("ent::k_binit_zero", "ent::k_binit_src", "ent::k_blevel", "ent::k_blevel_bu"), "ent::k_binit_zero"),
}


def _run_pmc(counter, out_dir, bench_args, timeout_s):
    d = os.path.join(out_dir, counter.lower() + "_" + bench_args[0].lstrip("-").replace("=", "_"))
    cmd = ["timeout", "-s", "KILL", str(timeout_s), "rocprofv3", "--pmc", counter, "-f", "csv", "-d", d,
           "-o", "run", "--", sys.executable, os.path.join(ROOT, "bench.py")] + list(bench_args)
    print("[traffic] " + " ".join(cmd), file=sys.stderr, flush=True)
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, cwd="/tmp")
    if r.returncode != 0:
        raise RuntimeError(f"{counter} run failed ({r.returncode}): {r.stderr.decode()[-2000:]}")
    lines = [x for x in r.stdout.decode().splitlines() if x.startswith("{")]
    bench_out = json.loads(lines[-1]) if lines else None
    path = None
    for dp, _, files in os.walk(d):
        for f in files:
            if f.endswith("counter_collection.csv"):
                path = os.path.join(dp, f)
    if path is None:
        raise RuntimeError(f"{counter}: no counter_collection.csv under {d}")
    return read_counters(path, counter), bench_out, " ".join(cmd)


def read_counters(path, counter):
    """[(kernel name, value in KiB)] for every dispatch of `counter`."""
    out = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                out.append((row["Kernel_Name"], float(row["Counter_Value"])))
    return out


def kernel_base(name):
    """'void ent::k_level<256>(ent::Args, long long)' -> 'ent::k_level'."""
    base = name.split("(")[0].split("<")[0].strip()
    return base[5:] if base.startswith("void ") else base


def per_search(rows, mode):
    kernels, starter = SEARCH_KERNELS[mode]
    tot, searches, by_kernel = 0.0, 0, {}
    for name, v in rows:
        base = kernel_base(name)
        if base == starter:
            searches += 1
        if base in kernels:
            tot += v
            by_kernel[base] = by_kernel.get(base, 0.0) + v
    return tot, searches, by_kernel


def calib_factor(rows, nbytes):
    vals = [v for name, v in rows if "k_calib_copy" in name]
    if not vals:
        raise RuntimeError("no k_calib_copy dispatch in the calibration run")
    vals.sort()
    med = vals[len(vals) // 2]
    return nbytes / (med * 1024.0), med


def measure(mode, bench_args, out_dir, timeout_s=240):
    """Runs the four counter passes; returns the summary dict."""
    os.makedirs(out_dir, exist_ok=True)
    nbytes = 1 << 30
    cf_rows, _, cf_cmd = _run_pmc("FETCH_SIZE", out_dir, ["--mode", "calib"], timeout_s)
    cw_rows, _, cw_cmd = _run_pmc("WRITE_SIZE", out_dir, ["--mode", "calib"], timeout_s)
    f_scale, f_copy = calib_factor(cf_rows, nbytes)
    w_scale, w_copy = calib_factor(cw_rows, nbytes)
    f_rows, f_out, f_cmd = _run_pmc("FETCH_SIZE", out_dir, bench_args, timeout_s)
    w_rows, w_out, w_cmd = _run_pmc("WRITE_SIZE", out_dir, bench_args, timeout_s)
    f_tot, f_n, f_by = per_search(f_rows, mode)
    w_tot, w_n, w_by = per_search(w_rows, mode)
    if f_n == 0 or w_n == 0:
        raise RuntimeError("no search kernels found in the counter runs")
    fetch_b = f_tot * 1024.0 * f_scale / f_n
    write_b = w_tot * 1024.0 * w_scale / w_n
    alg = None
    for o in (f_out, w_out):
        if o is not None:
            alg = o.get("roofline", {}).get("alg_bytes_per_bfs_mean") or o.get("roofline", {}).get("alg_bytes")
            break
    return {
        "mode": mode,
        "bench_args": list(bench_args),
        "searches": {"fetch_run": f_n, "write_run": w_n},
        "copy_kernel": {"bytes_read": nbytes, "bytes_written": nbytes, "FETCH_SIZE_KiB": f_copy,
                        "WRITE_SIZE_KiB": w_copy, "fetch_scale": round(f_scale, 4),
                        "write_scale": round(w_scale, 4)},
        "fetch_bytes_per_search": int(fetch_b),
        "write_bytes_per_search": int(write_b),
        "traffic_bytes_per_search": int(fetch_b + write_b),
        "alg_bytes_per_search": alg,
        "traffic_over_alg": round((fetch_b + write_b) / alg, 3) if alg else None,
        "fetch_KiB_by_kernel_per_search": {k: round(v / f_n, 1) for k, v in f_by.items()},
        "write_KiB_by_kernel_per_search": {k: round(v / w_n, 1) for k, v in w_by.items()},
        "commands": [cf_cmd, cw_cmd, f_cmd, w_cmd],
    }


def figure(name):
    """HBM bytes per search from the committed counter summary
    profiles/traffic_<name>.json, or None when it is absent."""
    t = load(os.path.join(ROOT, "profiles", f"traffic_{name}.json"))
    return None if t is None else t.get("traffic_bytes_per_search")


def load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None
