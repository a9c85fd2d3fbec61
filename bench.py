"""Benchmark of the MI355X Richards step on the 26.2 M-cell workload.

Workload richards_hetero_512x512x100 (512 x 512 x 100 cells, terrain-following,
log-normal Ks, 300 s steps): W untimed warmup steps, then K timed steps, for
the plain FP64 form and the mixed form (FP32 inner PCG inside FP64
refinement), each from the same initial state.  Inputs are resident in HBM
before the timed region starts.

metric      cell-iterations/s = active cells x PCG iterations / PCG wall time
            (the FP64 form is `value`; the mixed form is reported beside it)
roofline    the dominant kernel is the PCG direction kernel (p = z + beta p,
            q = A p): algorithmic bytes per launch / its mean launch time,
            measured with HIP events on the solver's stream
cpu_baseline  the FP64 oracle (C++/OpenMP, z-line PCG) on a bounded sample
            of the same workload: a 128 x 128 x 100 slice, 2 steps

One GPU only: --gpus N > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_form(case, steps, warmup, precision):
    from frehg2_amd.gw import RichardsSolver
    case.precision = precision
    t0 = time.perf_counter()
    g = RichardsSolver(case)
    g.synchronize()
    log(f"{precision}: solver staged in {time.perf_counter() - t0:.1f} s "
        f"({g.active_cells()} active cells)")
    for _ in range(warmup):
        g.step(case.dt)
    g.synchronize()
    g.set_timing(True)
    t0 = time.perf_counter()
    its = 0
    refs = 0
    for s in range(steps):
        g.step(case.dt)
        st = g.last_solve()
        its += st["iterations"]
        refs += st["refinements"]
    g.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing()
    n = g.active_cells()
    res = {
        "precision": precision,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "step_ms_events": tm["step_ms"] / max(tm["steps"], 1),
        "solve_ms_per_step": tm["solve_ms"] / max(tm["steps"], 1),
        "pcg_iterations": its,
        "iterations_per_solve": its / steps,
        "refinements_per_solve": refs / steps,
        "cell_iters_per_s": n * its / (tm["solve_ms"] * 1e-3),
        "cell_steps_per_s": n * steps / wall,
        "spmv_avg_ms": tm["spmv_ms"] / max(tm["spmv_launches"], 1),
        "spmv_launches": tm["spmv_launches"],
        "spmv_bytes": tm["spmv_bytes"],
        "active_cells": n,
    }
    res["spmv_gbs"] = res["spmv_bytes"] / (res["spmv_avg_ms"] * 1e-3) / 1e9
    from frehg2_amd import abi
    head = g.get(abi.F_HEAD)[1:-1, 1:-1]
    wc = g.get(abi.F_WC)[1:-1, 1:-1]
    log(f"{precision}: {res['ms_per_step']:.2f} ms/step, {res['iterations_per_solve']:.1f} it/solve, "
        f"{res['cell_iters_per_s'] / 1e9:.2f} G cell-it/s, spmv {res['spmv_avg_ms']:.3f} ms "
        f"({res['spmv_gbs']:.0f} GB/s)")
    g.close()
    return res, head, wc


def cpu_baseline():
    """FP64 oracle on a 128 x 128 x 100 slice of the same workload, 2 steps."""
    from frehg2_amd import cases
    from oracle import oracle
    threads = min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    c = cases.hetero_512x512x100(nx=128, ny=128, nz=100)
    o = oracle.OracleRichards(c)
    steps = 2
    for _ in range(steps):
        o.step(c.dt)
    t = o.timing()
    n = o.active_cells()
    return {
        "value": n * t["iterations"] / t["solve_s"],
        "unit": "cell-iterations/s",
        "cores": threads,
        "kind": "port",
        "sample": f"hetero workload generator at 128x128x100 ({n} cells), {steps} steps of "
                  f"300 s, z-line PCG rtol 1e-10; solve {t['solve_s']:.2f} s, "
                  f"{t['iterations']} iterations; whole steps {n * steps / t['step_s']:.3e} "
                  f"cell-steps/s",
    }


def load_traffic():
    p = ROOT / "profiles" / "traffic_pcg_dir_hetero_512x512x100.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              f"sharding of this path is described in DESIGN.md and not measured here.",
              file=sys.stderr)
        sys.exit(2)
    from frehg2_amd import cases
    t0 = time.perf_counter()
    case = cases.hetero_512x512x100(nx=args.nx, ny=args.ny, nz=args.nz)
    log(f"synthetic inputs generated in {time.perf_counter() - t0:.1f} s")
    forms = [f for f in args.forms.split(",") if f and not (args.no_mixed and f == "mixed")]
    results = {}
    heads = {}
    for f in forms:
        results[f], heads[f], _ = run_form(case, args.steps, args.warmup, f)
    r64 = results.get("fp64") or results[forms[0]]
    out = {
        "metric": "Richards cell-iters/s @26M cells: FP64 vs mixed-prec @1e-8 rel err; % HBM peak",
        "value": r64["cell_iters_per_s"],
        "unit": "cell-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if r64["precision"] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": case.name, "grid": [args.nx, args.ny, args.nz],
                   "cells": r64["active_cells"], "dt_s": case.dt,
                   "preconditioner": case.preconditioner, "rtol": case.rtol,
                   "parallelism": "single GPU"},
    }
    bw = r64["spmv_gbs"]
    out["roofline"] = {"bound": "hbm", "achieved": bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": bw / HBM_PEAK_GBS, "traffic": load_traffic(),
                       "kernel": "k_pcg_dir (FP64)", "avg_launch_ms": r64["spmv_avg_ms"],
                       "algorithmic_bytes_per_launch": r64["spmv_bytes"]}
    out["fp64"] = r64
    if "mixed" in results:
        rm = results["mixed"]
        hm, h64 = heads["mixed"], heads.get("fp64")
        rm["head_rel_err_vs_fp64"] = (float(np.max(np.abs(hm - h64)) / np.max(np.abs(h64)))
                                      if h64 is not None else None)
        rm["solve_speedup_vs_fp64"] = r64["solve_ms_per_step"] / rm["solve_ms_per_step"]
        rm["step_speedup_vs_fp64"] = r64["ms_per_step"] / rm["ms_per_step"]
        out["mixed"] = rm
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline()
        log(f"cpu baseline in {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
