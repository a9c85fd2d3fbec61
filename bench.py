"""Benchmark of the MI355X Richards step on the 26.2 M-cell workload.

Workload richards_hetero_512x512x100 (512 x 512 x 100 cells, terrain-following,
log-normal Ks, 300 s steps): W untimed warmup steps, then K timed steps, for
the plain FP64 form and the mixed form (FP32 inner PCG with FP64 residual
updates), each from the same initial state.  Inputs are resident in HBM
before the timed region starts.

metric      cell-iterations/s = active cells x PCG iterations / PCG wall time
            (the FP64 form is `value`; the mixed form is reported beside it)
accuracy    "@1e-8 rel err": each form runs at the tolerance that keeps its
            head within 1e-8 (max-abs, relative to max |h|) of a converged
            solution.  After the timed steps both forms take one step from the
            same state and are compared with the FP64 oracle's step from that
            state converged to rtol 1e-13; the errors are reported.
roofline    the dominant kernel is the FP64 PCG direction kernel (k_pcg_dir_p2,
            column pairs in the row-pair layout): algorithmic bytes per launch /
            its mean launch time, measured with HIP events on the solver's own
            stream over the timed steps, one launch in 8 (launches that exit at
            convergence are not counted); traffic = its HBM bytes per launch from
            two rocprofv3 counter runs (bench.py --write-traffic)
stream      frehg_hbm_probe, run first: the GB/s of an FP64 copy and of a
            6-read/2-write FP64 stream over eight fields of the workload's size
            (roofline.measured_stream_gbs, roofline.frac_of_measured_stream)
cpu_baseline  the FP64 oracle (C++/OpenMP, z-line PCG) on a bounded sample of
            the same workload: a 256 x 256 x 100 slice, 5 steps

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
ERR_TARGET = 1e-8
# One SpMV launch in SPMV_SAMPLE is bracketed by HIP events: an event pair
# delays the next dispatch by ~5 us, which over every launch would add ~0.45 ms
# to a step.
SPMV_SAMPLE = 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=100)
    # Matched accuracy: on this workload FP64 at rtol 1.5e-12 and the mixed
    # form at 3e-12 (FP32 residual reduction 1e-5 between FP64 updates) both
    # land near 5e-9 head error; FP64 at 3e-12 does not meet 1e-8.
    ap.add_argument("--rtol-fp64", type=float, default=1.5e-12,
                    help="FP64 form tolerance (meets 1e-8 head error on this workload)")
    ap.add_argument("--rtol-mixed", type=float, default=3e-12,
                    help="mixed form tolerance (meets 1e-8 head error on this workload)")
    ap.add_argument("--inner-rtol", type=float, default=1e-5,
                    help="mixed form: FP32 residual reduction between FP64 updates (0: library default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the accuracy check")
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the PCG launch shapes by timing trial solves before warmup")
    ap.add_argument("--pcg-shapes", default="",
                    help="PCG launch shapes as waves per CU: MARCH64,MARCH32,COLUMN (0 keeps)")
    ap.add_argument("--ensemble", action="store_true",
                    help="run the coupled-ensemble workload instead (64 members of 128x128x64)")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--member-shape", default="128,128,64")
    ap.add_argument("--columns", action="store_true",
                    help="run the uneven-columns workload instead (4096 columns of 50-400 cells)")
    ap.add_argument("--step-mode", default="host", choices=["host", "device", "graph"],
                    help="ensemble: step mode of the stacked legs (frehg_gw_set_step_mode)")
    ap.add_argument("--pc-sweep", action="store_true",
                    help="mixed form's inner preconditioners (zline, jacobi, chebyshev 3/5/7) on "
                         "the stiff 32x32x50 and the 512x512x100 workloads: time, iterations "
                         "and error against the FP64 oracle")
    ap.add_argument("--ens-rtol", type=float, default=0.0,
                    help="ensemble: PCG rtol of every leg (0: the workload's 1e-10)")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the direction kernel's HBM traffic JSON (roofline.traffic) "
                         "from the counter_collection.csv files of two rocprofv3 runs "
                         "(--pmc FETCH_SIZE, then --pmc WRITE_SIZE) of this bench, and print "
                         "every PCG kernel's traffic")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_case(args, precision, rtol):
    from frehg2_amd import cases
    c = cases.hetero_512x512x100(nx=args.nx, ny=args.ny, nz=args.nz, precision=precision)
    c.rtol = rtol
    if precision == "mixed" and args.inner_rtol > 0:
        c.inner_rtol = args.inner_rtol
    return c


def run_form(args, precision, rtol):
    from frehg2_amd.gw import RichardsSolver
    case = make_case(args, precision, rtol)
    t0 = time.perf_counter()
    g = RichardsSolver(case)
    g.synchronize()
    log(f"{precision}: solver staged in {time.perf_counter() - t0:.1f} s "
        f"({g.active_cells()} active cells), rtol {rtol:g}")
    tuned = None
    if args.pcg_shapes:
        g.set_pcg_shapes(*(int(v) for v in args.pcg_shapes.split(",")))
    if args.autotune:
        from frehg2_amd import autotune
        t1 = time.perf_counter()
        tuned = autotune.tune(g, dt=case.dt)
        log(f"{precision}: autotuned PCG shapes {tuned['shapes']} "
            f"({len(tuned['trials_ms'])} trials, {time.perf_counter() - t1:.1f} s)")
    for _ in range(args.warmup):
        g.step(case.dt)
    g.synchronize()
    g.set_timing(SPMV_SAMPLE)
    t0 = time.perf_counter()
    its = 0
    refs = 0
    for _ in range(args.steps):
        g.step(case.dt)
        st = g.last_solve()
        its += st["iterations"]
        refs += st["refinements"]
    g.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing()
    g.set_timing(False)
    n = g.active_cells()
    res = {
        "precision": precision,
        "rtol": rtol,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / args.steps,
        "step_ms_events": tm["step_ms"] / max(tm["steps"], 1),
        "solve_ms_per_step": tm["solve_ms"] / max(tm["steps"], 1),
        "pcg_iterations": its,
        "iterations_per_solve": its / args.steps,
        "refinements_per_solve": refs / args.steps,
        "cell_iters_per_s": n * its / (tm["solve_ms"] * 1e-3),
        "cell_steps_per_s": n * args.steps / wall,
        "spmv_avg_ms": tm["spmv_ms"] / max(tm["spmv_launches"], 1),
        "spmv_launches": tm["spmv_launches"],
        "spmv_bytes": tm["spmv_bytes"],
        "active_cells": n,
        "pcg_shapes_waves_per_cu": g.pcg_shapes(),
    }
    if tuned is not None:
        res["autotune_trials_solve_ms"] = tuned["trials_ms"]
    res["spmv_gbs"] = res["spmv_bytes"] / (res["spmv_avg_ms"] * 1e-3) / 1e9
    log(f"{precision}: {res['ms_per_step']:.2f} ms/step, solve {res['solve_ms_per_step']:.2f} ms, "
        f"{res['iterations_per_solve']:.1f} it/solve, {res['cell_iters_per_s'] / 1e9:.2f} G cell-it/s, "
        f"dir kernel {res['spmv_avg_ms']:.3f} ms ({res['spmv_gbs']:.0f} GB/s)")
    return res, g, case


def accuracy(args, solvers):
    """One step of each form from the FP64 form's final state, against the
    FP64 oracle's step from the same state (rtol 1e-13, the host's threads):
    head and theta error normwise, and each side's mass-audit residual
    dV - (boundary_in - ss_storage + realloc_adjust - vloss) relative to V."""
    from frehg2_amd import abi
    from oracle import oracle as orc
    threads = min(os.cpu_count() or 1, 16)
    orc.set_threads(threads)
    g64, case = solvers["fp64"]
    state = (g64.get(abi.F_HEAD), g64.get(abi.F_WC))

    def audit_residual(sim, v0):
        a = sim.audit()
        v = sim.volume()
        return ((v - v0) - (a["boundary_in"] - a["ss_storage"] + a["realloc_adjust"]
                            - a["vloss"])) / v

    t0 = time.perf_counter()
    ocase = make_case(args, "fp64", 1e-13)
    ocase.atol = 1e-300
    o = orc.OracleRichards(ocase)
    o.set(abi.F_HEAD, state[0])
    o.set(abi.F_WC, state[1])
    v0 = o.volume()
    o.step(case.dt)
    hT, wT = o.get(abi.F_HEAD)[1:-1, 1:-1], o.get(abi.F_WC)[1:-1, 1:-1]
    sT = o.last_solve()
    out = {"reference": (f"the FP64 oracle's step from the FP64 form's final state, rtol 1e-13 "
                         f"(residual {sT['residual_norm']:.3e}, ||b|| {sT['rhs_norm']:.3e}), "
                         f"{threads} host threads, {time.perf_counter() - t0:.1f} s"),
           "oracle_audit_residual": audit_residual(o, v0)}
    o.close()
    del o
    for f in solvers:
        g = solvers[f][0]
        g.set(abi.F_HEAD, state[0])
        g.set(abi.F_WC, state[1])
        v0 = g.volume()
        g.step(case.dt)
        h, w = g.get(abi.F_HEAD)[1:-1, 1:-1], g.get(abi.F_WC)[1:-1, 1:-1]
        out[f"{f}_head_rel_err"] = float(np.max(np.abs(h - hT)) / np.max(np.abs(hT)))
        out[f"{f}_theta_rel_err"] = float(np.max(np.abs(w - wT)) / np.max(np.abs(wT)))
        out[f"{f}_residual"] = g.last_solve()["residual_norm"]
        out[f"{f}_audit_residual"] = audit_residual(g, v0)
    out["target"] = ERR_TARGET
    out["met"] = all(out[f"{f}_head_rel_err"] <= ERR_TARGET for f in solvers)
    for f in solvers:
        log(f"accuracy {f} vs the FP64 oracle: head {out[f'{f}_head_rel_err']:.3e}, "
            f"theta {out[f'{f}_theta_rel_err']:.3e}, audit residual "
            f"{out[f'{f}_audit_residual']:.1e} (oracle {out['oracle_audit_residual']:.1e})")
    return out


def cpu_baseline():
    """FP64 oracle on a 256 x 256 x 100 slice of the same workload, 5 steps."""
    from frehg2_amd import cases
    from oracle import oracle
    threads = min(os.cpu_count() or 1, 16)
    oracle.set_threads(threads)
    c = cases.hetero_512x512x100(nx=256, ny=256, nz=100)
    c.rtol = 1e-12
    o = oracle.OracleRichards(c)
    steps = 5
    for _ in range(steps):
        o.step(c.dt)
    t = o.timing()
    n = o.active_cells()
    cpu = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    # The reference's default algebra (CG + ICC(0), triangular solves serial)
    # on a smaller slice, reported beside the port's z-line lane.
    ci = cases.hetero_512x512x100(nx=128, ny=128, nz=100)
    ci.rtol = 1e-12
    ci.preconditioner = "icc0"
    oi = oracle.OracleRichards(ci)
    steps_i = 2
    for _ in range(steps_i):
        oi.step(ci.dt)
    ti = oi.timing()
    ni = oi.active_cells()
    return {
        "value": n * t["iterations"] / t["solve_s"],
        "unit": "cell-iterations/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle (C++/OpenMP, z-line PCG, rtol 1e-12) on the same generator at "
                  f"256x256x100 ({n} cells), {steps} steps of 300 s: solve {t['solve_s']:.2f} s, "
                  f"{t['iterations']} iterations; whole steps {t['step_s']:.2f} s = "
                  f"{n * steps / t['step_s']:.3e} cell-steps/s; {threads} threads on {cpu}",
        "reference_algebra": {
            "value": ni * ti["iterations"] / ti["solve_s"],
            "unit": "cell-iterations/s",
            "sample": f"the reference's default algebra, CG + ICC(0) (triangular solves serial, the "
                      f"rest on {threads} threads), at 128x128x100 ({ni} cells), {steps_i} steps: "
                      f"solve {ti['solve_s']:.2f} s, {ti['iterations']} iterations; whole steps "
                      f"{ti['step_s']:.2f} s = {ni * steps_i / ti['step_s']:.3e} cell-steps/s",
        },
    }


def ensemble_accuracy(ens, gm, g64, solos, dt):
    """One step from the same state (the stacked mixed run's final state) by
    the stacked mixed solve, the stacked FP64 solve and each member's own FP64
    solve, against that step converged to rtol 1e-13 (stacked, FP64).  Per
    member: max |h - h_truth| / max |h_truth| and the error of the member's
    volume change over the step relative to that change."""
    from frehg2_amd import abi
    from frehg2_amd.gw import RichardsSolver
    names = ("eta", "dept", "seep_accum", "avail", "gain")
    state = {"h": gm.get(abi.F_HEAD), "w": gm.get(abi.F_WC)}
    cpl = {n: gm.coupling_state(n) for n in names}
    c0 = ens.members[0]
    cellv = c0.dx * c0.dy * c0.dz

    def load(g, h, w, cp):
        g.set(abi.F_HEAD, h)
        g.set(abi.F_WC, w)
        for n in names:
            g.set_coupling_state(n, cp[n])

    def member_state(m):
        h = np.zeros((ens.ny + 2, c0.nx + 2, c0.nz))
        w = np.zeros_like(h)
        h[1:-1, 1:-1] = ens.member_view(state["h"], m)
        w[1:-1, 1:-1] = ens.member_view(state["w"], m)
        return h, w, {n: ens.member_coupling(cpl[n], m) for n in names}

    ens.case.precision = "fp64"
    ens.case.rtol = 1e-13
    truth = RichardsSolver(ens.case)
    heads = {}
    closure = {}

    def closed_step(g):
        v0 = g.volume()
        g.step(dt)
        a = g.audit()
        dv = g.volume() - v0
        return a, (dv - (a["boundary_in"] - a["ss_storage"] + a["realloc_adjust"] - a["vloss"])) / abs(v0)

    for name, g in (("truth", truth), ("stacked_mixed", gm), ("stacked_fp64", g64)):
        load(g, state["h"], state["w"], cpl)
        a, closure[name] = closed_step(g)
        heads[name] = (g.get(abi.F_HEAD), g.get(abi.F_WC))
        if name == "stacked_mixed":
            audit = a
    truth.close()
    hT, wT = heads["truth"]
    out = {"reference": "one step of dt from the stacked mixed run's final state, stacked FP64 "
                        "converged to rtol 1e-13",
           "rtol_workload": c0.rtol}
    err = {k: {"head": [], "volume": []} for k in ("stacked_mixed", "stacked_fp64",
                                                    "sequential_fp64")}
    solo_audit = {k: 0.0 for k in audit}
    solo_closure = 0.0
    for m, r in enumerate(solos):
        h, w, cp = member_state(m)
        load(r, h, w, cp)
        a, res = closed_step(r)
        solo_closure = max(solo_closure, abs(res))
        for k in solo_audit:
            solo_audit[k] += a[k]
        v0 = float(np.sum(ens.member_view(state["w"], m))) * cellv
        dvT = float(np.sum(ens.member_view(wT, m))) * cellv - v0
        ht = ens.member_view(hT, m)
        sc = float(np.max(np.abs(ht)))
        got = {"stacked_mixed": tuple(ens.member_view(a, m) for a in heads["stacked_mixed"]),
               "stacked_fp64": tuple(ens.member_view(a, m) for a in heads["stacked_fp64"]),
               "sequential_fp64": (r.get(abi.F_HEAD)[1:-1, 1:-1], r.get(abi.F_WC)[1:-1, 1:-1])}
        for k, (hh, ww) in got.items():
            err[k]["head"].append(float(np.max(np.abs(hh - ht))) / sc)
            dv = float(np.sum(ww)) * cellv - v0
            err[k]["volume"].append(abs(dv - dvT) / max(abs(dvT), 1e-300))
    for k, e in err.items():
        out[k] = {"member_head_rel_err_max": max(e["head"]),
                  "member_head_rel_err_median": float(np.median(e["head"])),
                  "member_volume_change_rel_err_max": max(e["volume"]),
                  "member_volume_change_rel_err_median": float(np.median(e["volume"]))}
        log(f"ensemble accuracy {k}: head max {max(e['head']):.2e}, "
            f"volume change max {max(e['volume']):.2e}")
    out["audit_closure_rel"] = {k: abs(v) for k, v in closure.items()}
    out["audit_closure_rel"]["sequential_fp64_max_member"] = solo_closure
    scale = max(abs(v) for v in audit.values()) + 1e-300
    out["stacked_mixed_audit"] = audit
    out["stacked_mixed_audit_vs_sum_of_members_rel"] = max(
        abs(audit[k] - solo_audit[k]) for k in audit) / scale
    log(f"ensemble audit closure (relative to volume): " +
        ", ".join(f"{k} {v:.1e}" for k, v in out["audit_closure_rel"].items()))
    return out


def ensemble_bench(args):
    """Coupled ensemble: members stacked along y behind inactive separator
    rows, advanced by one solver (FP64 and mixed), against the same members
    run one after another by their own FP64 solvers at the same dt sequence.
    Every substep is a coupled step in lockstep (frehg2_amd.coupler.Coupler,
    surface at rest): depth refresh, window start, the groundwater step and
    the seepage application."""
    from frehg2_amd import abi, cases
    from frehg2_amd.coupler import Coupler, closure_residual
    from frehg2_amd.ensemble import Ensemble
    from frehg2_amd.gw import RichardsSolver
    nx, ny, nz = (int(v) for v in args.member_shape.split(","))
    M = args.members

    def bottom_of(c):
        return np.pad(Ensemble._bath(c), 1, mode="edge")

    def members(precision):
        out = []
        for m in range(M):
            c = cases.coupled_128x128x64(member=m, nx=nx, ny=ny, nz=nz, precision=precision)
            if args.ens_rtol > 0:
                c.rtol = args.ens_rtol
            if precision == "mixed" and args.inner_rtol > 0:
                c.inner_rtol = args.inner_rtol
            out.append(c)
        return out

    t0 = time.perf_counter()
    ens = Ensemble(members("mixed"))
    log(f"ensemble: {M} members of {nx}x{ny}x{nz} stacked to {ens.case.nx}x{ens.case.ny}x"
        f"{ens.case.nz} in {time.perf_counter() - t0:.1f} s")
    legs = {}
    bottom = ens.stack_plane([bottom_of(c) for c in ens.members])
    dt0 = ens.case.dt_init
    # mixed stacked leg: the adaptive controller picks the dt sequence that
    # every other leg replays
    ens.case.precision = "mixed"
    g = RichardsSolver(ens.case)
    g.set_step_mode(args.step_mode)
    cm = Coupler(g, bottom, dt=dt0, mode="sync")
    warm_dts, dts = [], []
    t = 0.0
    for _ in range(args.warmup):
        warm_dts.append(cm.next_dt())
        t += warm_dts[-1]
        cm.step(t, warm_dts[-1])
    g.synchronize()
    if args.step_mode == "host":
        g.set_timing(SPMV_SAMPLE)  # timing runs host-polled steps
    its = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dts.append(cm.next_dt())
        t += dts[-1]
        cm.step(t, dts[-1])
        its += g.last_solve()["iterations"]
    g.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing()
    legs["stacked_mixed"] = {"ms_per_substep": 1e3 * wall / args.steps,
                             "solve_ms_per_substep": tm["solve_ms"] / max(tm["steps"], 1),
                             "iterations_per_solve": its / args.steps,
                             "active_cells": g.active_cells(),
                             "cell_steps_per_s": g.active_cells() * args.steps / wall}
    gm = g
    log(f"ensemble stacked mixed: {legs['stacked_mixed']['ms_per_substep']:.2f} ms/substep, "
        f"{its / args.steps:.1f} it/solve")
    # the coupled closure identity (Coupler.hpp:45-56) over three more
    # substeps of this leg, outside the timed region
    vs, vg, acc = cm.surface_volume(), g.volume(), cm.audit().accum_volume
    vtot = vs + vg
    cpl_closure = 0.0
    for _ in range(3):
        dt = cm.next_dt()
        t += dt
        cm.step(t, dt)
        a = cm.audit()
        res = closure_residual(a, cm.surface_volume() - vs, g.volume() - vg,
                               a.accum_volume - acc)
        cpl_closure = max(cpl_closure, abs(res) / vtot)
        vs, vg, acc = cm.surface_volume(), g.volume(), a.accum_volume
    log(f"ensemble coupled closure (stacked mixed, relative to volume): {cpl_closure:.1e}")

    ens.case.precision = "fp64"
    g = RichardsSolver(ens.case)
    g.set_step_mode(args.step_mode)
    c64 = Coupler(g, bottom, dt=dt0, mode="sync")
    t = 0.0
    for dt in warm_dts:
        t += dt
        c64.step(t, dt)
    g.synchronize()
    if args.step_mode == "host":
        g.set_timing(SPMV_SAMPLE)
    its = 0
    t0 = time.perf_counter()
    for dt in dts:
        t += dt
        c64.step(t, dt)
        its += g.last_solve()["iterations"]
    g.synchronize()
    wall = time.perf_counter() - t0
    tm = g.timing()
    legs["stacked_fp64"] = {"ms_per_substep": 1e3 * wall / args.steps,
                            "solve_ms_per_substep": tm["solve_ms"] / max(tm["steps"], 1),
                            "iterations_per_solve": its / args.steps,
                            "cell_steps_per_s": g.active_cells() * args.steps / wall}
    g64 = g
    log(f"ensemble stacked fp64: {legs['stacked_fp64']['ms_per_substep']:.2f} ms/substep")

    solo_cases = members("fp64")
    solos = [RichardsSolver(c) for c in solo_cases]
    couplers = [Coupler(r, bottom_of(c), dt=dt0, mode="sync") for r, c in zip(solos, solo_cases)]
    t = 0.0
    for dt in warm_dts:
        t += dt
        for cp in couplers:
            cp.step(t, dt)
    for r in solos:
        r.synchronize()
    t0 = time.perf_counter()
    for dt in dts:
        t += dt
        for cp in couplers:
            cp.step(t, dt)
    for r in solos:
        r.synchronize()
    wall = time.perf_counter() - t0
    n = sum(r.active_cells() for r in solos)
    legs["sequential_fp64"] = {"ms_per_substep": 1e3 * wall / args.steps,
                               "cell_steps_per_s": n * args.steps / wall}
    log(f"ensemble sequential fp64: {legs['sequential_fp64']['ms_per_substep']:.2f} ms/substep")
    accuracy = ensemble_accuracy(ens, gm, g64, solos, dts[-1])
    for r in solos:
        r.close()
    gm.close()
    g64.close()
    seq = legs["sequential_fp64"]["ms_per_substep"]
    out = {
        "metric": "coupled ensemble substep throughput: stacked mixed vs sequential FP64 solves",
        "value": legs["stacked_mixed"]["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": legs["stacked_mixed"]["ms_per_substep"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": "richards_coupled_128x128x64 ensemble", "members": M,
                   "step_mode_stacked": args.step_mode,
                   "member_grid": [nx, ny, nz], "stacked_grid": [ens.case.nx, ens.case.ny, nz],
                   "dt": "adaptive (common to all members)",
                   "coupling": "sync coupled step per substep, surface at rest",
                   "parallelism": "single GPU"},
        "legs": legs,
        "speedup_stacked_mixed_vs_sequential_fp64": seq / legs["stacked_mixed"]["ms_per_substep"],
        "speedup_stacked_fp64_vs_sequential_fp64": seq / legs["stacked_fp64"]["ms_per_substep"],
        "accuracy": accuracy,
        "coupled_closure_rel_max": cpl_closure,
        "dt_sequence": dts,
    }
    print(json.dumps(out), flush=True)


def columns_bench(args):
    """4096 independent columns of 50-400 cells (lateral Ks = 0, beds lowered
    per column): FP64 and mixed forms, each stepped host-polled and as a
    replayed graph, from the same state; the adaptive controller's dt
    sequence of the first leg is replayed by the others."""
    from frehg2_amd import cases
    from frehg2_amd.gw import RichardsSolver
    legs = {}
    warm, dts = [], []
    for form, mode in (("fp64", "host"), ("fp64", "graph"), ("mixed", "host"),
                       ("mixed", "graph")):
        c = cases.columns_4096()
        c.precision = form
        c.rtol = 1e-10
        g = RichardsSolver(c)
        g.set_step_mode(mode)
        n = g.active_cells()
        first = not dts

        def one(s, timed):
            for which, vals in c.bc_schedule(s).items():
                g.set_bc_values(which, vals)
            if first:
                dt = g.next_dt()
                (dts if timed else warm).append(dt)
            else:
                dt = dts[s - args.warmup] if timed else warm[s]
            g.step(dt)
            return g.last_solve()["iterations"]

        for s in range(args.warmup):
            one(s, False)
        g.synchronize()
        its = 0
        t0 = time.perf_counter()
        for s in range(args.warmup, args.warmup + args.steps):
            its += one(s, True)
        g.synchronize()
        wall = time.perf_counter() - t0
        key = form if mode == "host" else f"{form}_{mode}"
        legs[key] = {"ms_per_step": 1e3 * wall / args.steps, "active_cells": n,
                     "iterations_per_solve": its / args.steps,
                     "cell_steps_per_s": n * args.steps / wall,
                     "column_steps_per_s": 4096 * args.steps / wall,
                     "step_mode": g.graph_stats()}
        log(f"columns {key}: {legs[key]['ms_per_step']:.3f} ms/step, "
            f"{legs[key]['iterations_per_solve']:.2f} it/solve, {n} active cells")
        g.close()
    best = min(legs, key=lambda k: legs[k]["ms_per_step"])
    out = {
        "metric": "uneven-columns step throughput (4096 columns of 50-400 cells)",
        "value": legs["fp64_graph"]["column_steps_per_s"],
        "unit": "column-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": legs["fp64_graph"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": "richards_columns_4096", "grid": [64, 64, 400],
                   "dt": "adaptive (first leg's sequence replayed)", "step_mode": "graph",
                   "parallelism": "single GPU"},
        "legs": legs,
        "fastest_leg": best,
        "graph_speedup_fp64": legs["fp64"]["ms_per_step"] / legs["fp64_graph"]["ms_per_step"],
        "graph_speedup_mixed": legs["mixed"]["ms_per_step"] / legs["mixed_graph"]["ms_per_step"],
        "mixed_speedup": legs["fp64_graph"]["ms_per_step"] / legs["mixed_graph"]["ms_per_step"],
    }
    print(json.dumps(out), flush=True)


def pc_sweep(args):
    """The mixed form's inner preconditioners on the stiff 32x32x50 workload
    (adaptive dt, 20 steps) and on the 512x512x100 one (dt 300 s, 5 steps):
    ms per step, PCG iterations per solve, and head / theta error of one
    more step against the FP64 oracle's step from the same state."""
    import dataclasses
    from frehg2_amd import abi, cases
    from frehg2_amd.gw import RichardsSolver
    from oracle import oracle as orc
    from oracle.oracle import OracleRichards
    orc.set_threads(min(os.cpu_count() or 1, 16))
    variants = [("zline", 0), ("jacobi", 0), ("chebyshev", 3), ("chebyshev", 5), ("chebyshev", 7)]

    def nerr(a, b):
        return float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))

    legs = {}
    for label, make, steps, dt_fixed, rtol in (
            ("richards_stiff_32x32x50", lambda: cases.stiff_32x32x50(precision="mixed"), 20,
             None, 1e-10),
            ("richards_hetero_512x512x100", lambda: cases.hetero_512x512x100(precision="mixed"),
             5, 300.0, args.rtol_mixed)):
        base = make()
        for pc, deg in variants:
            key = f"{label}:{pc}{deg or ''}"
            c = dataclasses.replace(base)
            c.preconditioner, c.precision, c.rtol = pc, "mixed", rtol
            c.inner_rtol = args.inner_rtol
            c.max_iterations = 10000
            g = RichardsSolver(c)
            if pc == "chebyshev":
                g.set_chebyshev(deg, 0.05)
            for _ in range(2):
                g.step(g.next_dt() if c.adaptive else dt_fixed)
            g.synchronize()
            dts, its = [], 0
            t0 = time.perf_counter()
            for _ in range(steps):
                dts.append(g.next_dt() if c.adaptive else dt_fixed)
                g.step(dts[-1])
                its += g.last_solve()["iterations"]
            g.synchronize()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            co = dataclasses.replace(base)
            co.preconditioner, co.precision, co.rtol = "zline", "fp64", 1e-13
            o = OracleRichards(co)
            o.set(abi.F_HEAD, g.get(abi.F_HEAD))
            o.set(abi.F_WC, g.get(abi.F_WC))
            o.step(dts[-1])
            g.step(dts[-1])
            legs[key] = {"ms_per_step": ms, "iterations_per_solve": its / steps,
                         "head_err_vs_oracle": nerr(g.get(abi.F_HEAD)[1:-1, 1:-1],
                                                    o.get(abi.F_HEAD)[1:-1, 1:-1]),
                         "theta_err_vs_oracle": nerr(g.get(abi.F_WC)[1:-1, 1:-1],
                                                     o.get(abi.F_WC)[1:-1, 1:-1])}
            log(f"{key}: {ms:.3f} ms/step, {its / steps:.1f} it/solve, head err "
                f"{legs[key]['head_err_vs_oracle']:.1e}")
            g.close()
    out = {"metric": "mixed-form inner preconditioners: ms per step", "value":
           legs["richards_hetero_512x512x100:zline"]["ms_per_step"], "unit": "ms/step",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f32+f64",
           "data": "synthetic (seeded generator)",
           "config": {"workload": "richards_stiff_32x32x50 and richards_hetero_512x512x100",
                      "chebyshev_interval": "[0.1, 2] of D^-1 A"},
           "legs": legs}
    print(json.dumps(out), flush=True)


TRAFFIC_JSON = ROOT / "profiles" / "traffic_pcg_dir_hetero_512x512x100.json"
# the FP64 direction kernel the solver runs at 512 x 512 x 100 (column pairs, row pairs)
DIR_KERNEL = "k_pcg_dir_p2<double, 1, false, true>"


def write_traffic(fetch_csv, write_csv, nx=512, ny=512, nz=100):
    """HBM bytes per dispatch of the PCG kernels from two rocprofv3 counter
    runs.  FETCH_SIZE and WRITE_SIZE (KiB a dispatch) are scaled by what they
    read for the step's device-to-device copy of one FP64 field of known size
    ((nx+2)(ny+2)nz doubles, __amd_rocclr_copyBuffer); a kernel's figure is the
    median over its dispatches that moved more than 1 % of its largest (the
    launches queued past convergence return at once).  Writes the direction
    kernel's bytes per launch to the JSON bench.py reads as roofline.traffic."""
    import csv
    field = (nx + 2) * (ny + 2) * nz * 8.0

    def per_kernel(path, counter):
        vals = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"].replace("void ", "").replace(
                    "frehg::(anonymous namespace)::", "").split("(")[0]
                vals.setdefault(name, []).append(float(r["Counter_Value"]) * 1024.0)
        return vals

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")

    def copy_factor(vals):
        big = [v for v in vals.get("__amd_rocclr_copyBuffer", []) if v > 0.25 * field]
        return field / float(np.median(big)) if big else None

    ff, fw = copy_factor(fetch), copy_factor(write)
    if ff is None or fw is None:
        raise SystemExit("write-traffic: no field-sized copy in the counter runs to calibrate")
    active = nx * ny * nz
    rows = {}
    for name in sorted(set(fetch) & set(write)):
        if not name.startswith("k_pcg") and not name.startswith("k_sweep"):
            continue
        fv, wv = np.array(fetch[name]), np.array(write[name])
        f_med = float(np.median(fv[fv > 0.01 * fv.max()])) * ff
        w_med = float(np.median(wv[wv > 0.01 * wv.max()])) * fw
        rows[name] = (f_med, w_med)
        print(f"{name:48s} read {f_med / 1e6:8.1f} MB  write {w_med / 1e6:8.1f} MB  "
              f"{(f_med + w_med) / active:6.1f} B/cell", flush=True)
    if DIR_KERNEL not in rows:
        raise SystemExit(f"write-traffic: {DIR_KERNEL} not in the counter runs")
    f_med, w_med = rows[DIR_KERNEL]
    out = {"kernel": DIR_KERNEL, "workload": "richards_hetero_512x512x100",
           "bytes_per_launch": f_med + w_med, "fetch_scaled": f_med, "write_scaled": w_med,
           "fetch_factor": ff, "write_factor": fw, "algorithmic_bytes": 64.0 * active}
    TRAFFIC_JSON.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out), flush=True)


def load_traffic():
    p = TRAFFIC_JSON
    if p.exists():
        try:
            t = json.loads(p.read_text())
            # only a figure measured on the kernel this build runs
            return t.get("bytes_per_launch") if t.get("kernel") == DIR_KERNEL else None
        except (OSError, ValueError):
            return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              f"sharding of this path is described in DESIGN.md and not measured here.",
              file=sys.stderr)
        sys.exit(2)
    if args.ensemble:
        ensemble_bench(args)
        return
    if args.columns:
        columns_bench(args)
        return
    if args.pc_sweep:
        pc_sweep(args)
        return
    if args.write_traffic:
        write_traffic(*args.write_traffic, nx=args.nx, ny=args.ny, nz=args.nz)
        return
    # the achievable streaming rate on this device, measured live: the best
    # copy and 6-read/2-write FP64 stream over access width, cache policy and
    # occupancy, on eight fields of this workload's size
    from frehg2_amd import hbm
    sw = hbm.sweep((args.nx + 2) * (args.ny + 2) * args.nz)
    stream = {"copy_gbs": sw["best"]["copy"]["gbs"], "stream_6r2w_gbs": sw["best"]["6r2w"]["gbs"],
              "stream_5r1w_gbs": sw["best"]["5r1w"]["gbs"],
              "best_configs": {k: {kk: v[kk] for kk in ("bytes_per_lane", "nontemporal",
                                                        "blocks_per_cu")}
                               for k, v in sw["best"].items()}}
    log(f"hbm probe (best of {len(sw['table'])} configurations): copy "
        f"{stream['copy_gbs']:.0f} GB/s, 6r2w stream {stream['stream_6r2w_gbs']:.0f} GB/s, "
        f"5r1w {stream['stream_5r1w_gbs']:.0f} GB/s")
    forms = [f for f in args.forms.split(",") if f]
    results = {}
    solvers = {}
    for f in forms:
        rtol = args.rtol_fp64 if f == "fp64" else args.rtol_mixed
        results[f], g, case = run_form(args, f, rtol)
        solvers[f] = (g, case)
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
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": "richards_hetero_512x512x100", "grid": [args.nx, args.ny, args.nz],
                   "cells": r64["active_cells"], "dt_s": 300.0, "preconditioner": "zline",
                   "rtol_fp64": args.rtol_fp64, "rtol_mixed": args.rtol_mixed,
                   "parallelism": "single GPU"},
    }
    bw = r64["spmv_gbs"]
    out["roofline"] = {"bound": "hbm", "achieved": bw, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": bw / HBM_PEAK_GBS, "traffic": load_traffic(),
                       "kernel": DIR_KERNEL + " (FP64 p = z + beta p, q = A p, p.q over "
                                              "column pairs in the row-pair layout)",
                       "avg_launch_ms": r64["spmv_avg_ms"],
                       "algorithmic_bytes_per_launch": r64["spmv_bytes"],
                       "measured_stream_gbs": {"copy": stream["copy_gbs"],
                                               "stream_6r2w": stream["stream_6r2w_gbs"],
                                               "stream_5r1w": stream["stream_5r1w_gbs"],
                                               "best_configs": stream["best_configs"]},
                       "frac_of_measured_stream": bw / stream["stream_6r2w_gbs"]}
    out["fp64"] = r64
    if "mixed" in results:
        rm = results["mixed"]
        rm["solve_speedup_vs_fp64"] = r64["solve_ms_per_step"] / rm["solve_ms_per_step"]
        rm["step_speedup_vs_fp64"] = r64["ms_per_step"] / rm["ms_per_step"]
        out["mixed"] = rm
    if not args.no_accuracy and "fp64" in solvers:
        out["accuracy"] = accuracy(args, solvers)
    for g, _ in solvers.values():
        g.close()
    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline()
        log(f"cpu baseline in {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
