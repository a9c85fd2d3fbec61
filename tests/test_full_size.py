"""The benchmark configurations at their full sizes on the HIP path.

* richards_hetero_512x512x100 (26.2 M cells): one step from the seeded
  initial state on the GPU and on the oracle (16 host threads), identical
  inputs; head and theta within 1e-10 normwise, fluxes, audit and volume as in
  the small-size parity gate.  The mixed form's step is held to 1e-8 of the
  oracle's (the "@1e-8 rel err vs the FP64 oracle" of the benchmark).
* richards_columns_4096 (4096 uneven columns, 919,455 cells): every column,
  identical-input steps against the oracle.
* richards_coupled_128x128x64 x 64 (67.1 M cells stacked): coupled substeps
  of the stacked mixed form through the coupler; the audit closes every
  substep, the coupled window closure holds, and two members take the same
  identical-input step in the stack as in standalone FP64 solvers.
"""
import os

import numpy as np
import pytest

from frehg2_amd import abi, cases
from helpers import interior

pytestmark = pytest.mark.gpu


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _nerr(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def test_hetero_512x512x100_step_matches_the_oracle():
    from test_gpu_parity import _compare
    from frehg2_amd.gw import RichardsSolver
    from oracle import oracle as orc
    orc.set_threads(_threads())
    case = cases.hetero_512x512x100()
    case.rtol, case.atol = 1e-12, 1e-300
    print(f"[full] {case.name} {case.nx}x{case.ny}x{case.nz}: oracle on {_threads()} threads",
          flush=True)
    o = orc.OracleRichards(case)
    g = RichardsSolver(case)
    print("[full] created; stepping", flush=True)
    # identical inputs: the oracle's initial state (its theta(h) is the pow()
    # form, the device's the log form: they agree to rounding)
    assert np.array_equal(g.get(abi.F_HEAD), o.get(abi.F_HEAD))
    assert _nerr(g.get(abi.F_WC), o.get(abi.F_WC)) <= 1e-15
    g.set(abi.F_HEAD, o.get(abi.F_HEAD))
    g.set(abi.F_WC, o.get(abi.F_WC))
    o.step(case.dt)
    print(f"[full] oracle step done: {o.last_solve()}", flush=True)
    g.step(case.dt)
    print(f"[full] gpu step done: {g.last_solve()}", flush=True)
    errs = _compare(o, g, "hetero full size")
    print("[full] fp64 vs oracle: " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()),
          flush=True)
    h_ref = interior(o.get(abi.F_HEAD))
    w_ref = interior(o.get(abi.F_WC))
    g.close()
    mc = cases.hetero_512x512x100()
    mc.precision = "mixed"
    mc.rtol, mc.atol = 3e-12, 1e-300
    m = RichardsSolver(mc)
    m.set(abi.F_HEAD, o.get(abi.F_HEAD_N))
    m.set(abi.F_WC, o.get(abi.F_WC_N))
    m.step(mc.dt)
    eh = _nerr(interior(m.get(abi.F_HEAD)), h_ref)
    ew = _nerr(interior(m.get(abi.F_WC)), w_ref)
    print(f"[full] mixed vs oracle: head {eh:.2e} theta {ew:.2e} {m.last_solve()}", flush=True)
    assert eh <= 1e-8 and ew <= 1e-8


def test_columns_4096_every_column_matches_the_oracle():
    from test_gpu_parity import _stepwise
    from oracle import oracle as orc
    orc.set_threads(_threads())
    case = cases.columns_4096()
    assert case.nx * case.ny == 4096
    o, g = _stepwise(case, 12)
    assert g.active_cells() == o.active_cells() == 919455


def test_coupled_ensemble_64_members_full_size():
    from frehg2_amd.coupler import Coupler, DeviceExchange, closure_residual
    from frehg2_amd.ensemble import Ensemble
    from frehg2_amd.gw import RichardsSolver
    members = [cases.coupled_128x128x64(member=m) for m in range(64)]
    for c in members:
        c.rtol, c.atol = 1e-13, 1e-300
    ens = Ensemble(members)
    print(f"[full] stacked ensemble {ens.case.nx}x{ens.case.ny}x{ens.case.nz} "
          f"({ens.case.precision})", flush=True)

    def bottom_of(c):
        return np.pad(Ensemble._bath(c), 1, mode="edge")

    stacked = RichardsSolver(ens.case)
    assert stacked.active_cells() == 64 * 128 * 128 * 64
    cm = Coupler(stacked, ens.stack_plane([bottom_of(c) for c in members]),
                 dt=ens.case.dt_init, mode="sync")
    # two members alone, in the stacked run's precision and tolerance, each
    # with its own window kernels
    solos = {}
    for m in (0, 63):
        c = cases.coupled_128x128x64(member=m)
        c.rtol, c.atol = 1e-13, 1e-300
        solos[m] = RichardsSolver(c)
    windows = {m: DeviceExchange(solos[m], bottom_of(members[m])) for m in solos}
    names = ("eta", "dept", "avail", "seep_accum", "gain")
    vs, vg, acc = cm.surface_volume(), stacked.volume(), cm.audit().accum_volume
    t = 0.0
    worst = 0.0
    for s in range(3):
        # identical inputs for the two standalone members
        h, w = stacked.get(abi.F_HEAD), stacked.get(abi.F_WC)
        for m, solo in solos.items():
            solo.set(abi.F_HEAD, _member_field(ens, h, m, solo))
            solo.set(abi.F_WC, _member_field(ens, w, m, solo))
            for n in names:
                solo.set_coupling_state(n, ens.member_coupling(stacked.coupling_state(n), m))
            # the window start the coupler runs before the stacked step: the
            # depth refresh and avail = dept, gain = 0 (Exchange.cpp:21-52)
            windows[m].refresh_depth()
            windows[m].begin_window()
        dt = cm.next_dt()
        t += dt
        cm.step(t, dt)
        for m, solo in solos.items():
            solo.step(dt)
            eh = _nerr(ens.member_view(stacked.get(abi.F_HEAD), m), interior(solo.get(abi.F_HEAD)))
            ew = _nerr(ens.member_view(stacked.get(abi.F_WC), m), interior(solo.get(abi.F_WC)))
            worst = max(worst, eh, ew)
            if not (eh <= 1e-8 and ew <= 1e-8):
                dw = np.abs(ens.member_view(stacked.get(abi.F_WC), m) -
                            interior(solo.get(abi.F_WC)))
                print(f"[full] member {m}: {int(np.sum(dw > 1e-8))} cells differ > 1e-8",
                      flush=True)
            assert eh <= 1e-8 and ew <= 1e-8, (s, m, eh, ew)
        v = stacked.volume()
        ca = cm.audit()
        closure = closure_residual(ca, cm.surface_volume() - vs, v - vg, ca.accum_volume - acc)
        vtot = vs + vg
        print(f"[full] substep {s}: dt {dt:.3g} window closure {closure / vtot:.2e} "
              f"member error {worst:.2e} solve {stacked.last_solve()}", flush=True)
        assert abs(closure) <= 1e-12 * vtot, (s, closure)
        vs, vg, acc = cm.surface_volume(), v, ca.accum_volume


def _member_field(ens, field, m, solo):
    out = np.zeros(solo.field_shape(abi.F_HEAD))
    out[1:-1, 1:-1] = ens.member_view(field, m)
    # ghosts are rebuilt by set_field from the interior
    return out
