"""Device-controlled and graph-captured steps (frehg_gw_set_step_mode).

The reference's coupled substep (src/coupling/Coupler.cpp:106-119) runs
RichardsSolver::step, whose KSP solve (src/core/LinearSystem.cpp:320)
iterates under host control.  DEVICE mode takes the host-polled solve's
decisions in one-block control kernels on the device, so a step is a fixed
launch sequence; GRAPH captures that sequence once per buffer parity and
replays it, the step's dt read from a pinned host word.

Pins (every one bitwise: the same launches must give the same bits):
* GRAPH replays are the DEVICE step, step after step, with adaptive dt (each
  replay runs a new dt) and eager reads and uploads between replays, on the
  stacked coupled ensemble (FP64 and mixed) and on the columns workload with
  its per-step rain schedule;
* with a budget too small for the solves the gate closes and the host
  finishes the step: still bitwise the DEVICE step, which closes it too;
* a budget the solves fit in: DEVICE is HOST, FP64 and mixed (the mixed
  form's FP32 scale, 1 / sqrt(rr / n), and tolerances are computed on the
  device, whose double sqrt and divide round as the host's), also on a
  100-row terrain grid whose PCG runs the pair and quad kernels;
* a configuration change after capture (a sink) is captured again.
"""
import numpy as np
import pytest

from frehg2_amd import abi, cases
from frehg2_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

FIELDS = (abi.F_HEAD, abi.F_WC, abi.F_QX, abi.F_QY, abi.F_QZF, abi.F_VLOSS)
CPL = ("eta", "dept", "avail", "seep_accum", "gain")


def _stacked(precision="fp64", rtol=1e-12):
    ms = []
    for m in range(3):
        c = cases.coupled_128x128x64(member=m, nx=10, ny=7, nz=16, precision=precision)
        c.rtol, c.atol, c.max_iterations = rtol, 1e-300, 5000
        ms.append(c)
    return Ensemble(ms).case


def _terrain(precision="fp64"):
    """nz = 100, nx even: the PCG's FP64 column-pair and row-pair kernels and
    sweeps, and the FP32 quad kernels (the stacked members' 16-row columns
    take the one-row and row-pair forms)."""
    c = cases.hetero_512x512x100(nx=16, ny=10, nz=100, precision=precision)
    c.rtol, c.atol = (1e-12 if precision == "fp64" else 3e-12), 1e-300
    return c


def _columns(precision="fp64"):
    c = cases.columns_4096(ncol_side=12, lmax=150)
    c.precision = precision
    return c


def _run(case, modes, steps, budget=(0, 0), every=None):
    from frehg2_amd.gw import RichardsSolver
    sims = [RichardsSolver(case) for _ in modes]
    for s, m in zip(sims, modes):
        s.set_step_mode(m, *budget)
    for k in range(steps):
        if case.bc_schedule is not None:
            for which, vals in case.bc_schedule(k).items():
                for s in sims:
                    s.set_bc_values(which, vals)
        dts = [s.next_dt() for s in sims]
        assert len(set(dts)) == 1, (k, dts)
        for s in sims:
            s.step(dts[0])
        if every is not None:
            every(k, sims)
    return sims


def _same(a, b, k, coupled=False):
    for f in FIELDS:
        x, y = a.get(f), b.get(f)
        assert np.array_equal(x, y), (k, f, float(np.max(np.abs(x - y))))
    assert a.audit() == b.audit(), k
    assert a.last_solve() == b.last_solve(), k
    if coupled:
        for n in CPL:
            assert np.array_equal(a.coupling_state(n), b.coupling_state(n)), (k, n)


@pytest.mark.parametrize("name,make,coupled", [
    ("stacked_fp64", lambda: _stacked("fp64"), True),
    ("stacked_mixed", lambda: _stacked("mixed", 1e-10), True),
    ("columns_fp64", lambda: _columns("fp64"), False),
    ("terrain_fp64", lambda: _terrain("fp64"), False),
    ("terrain_mixed", lambda: _terrain("mixed"), False),
])
def test_graph_replay_is_the_device_step(name, make, coupled):
    case = make()
    print(f"[graph] {name} ({case.nx},{case.ny},{case.nz})", flush=True)
    sims = _run(case, ("device", "graph"), 8,
                every=lambda k, s: _same(s[0], s[1], k, coupled))
    st = sims[1].graph_stats()
    print(f"[graph] {name}: {st}", flush=True)
    assert st["replays"] == 7 and st["captures"] >= 2  # the first step sizes the budget
    assert sims[0].graph_stats()["replays"] == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_closes_the_gate_like_the_device_step(precision):
    """A budget of 1 launch pair in one round: most solves outrun it, the
    step's later kernels return at once and the host finishes the step."""
    case = _stacked(precision, 1e-12 if precision == "fp64" else 1e-10)
    print(f"[graph] tight budget {precision}", flush=True)
    sims = _run(case, ("device", "graph"), 5, budget=(1, 1),
                every=lambda k, s: _same(s[0], s[1], k, True))
    finishes = [s.graph_stats()["host_finishes"] for s in sims]
    print(f"[graph] tight budget {precision}: host finishes {finishes}", flush=True)
    assert finishes[0] == finishes[1] >= 3
    for s in sims:
        assert s.graph_stats()["budget_iterations"] == 1
        assert s.last_solve()["converged"] == 1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_step_is_the_host_step(precision):
    case = _stacked(precision, 1e-12 if precision == "fp64" else 1e-10)
    sims = _run(case, ("host", "device"), 6,
                every=lambda k, s: _same(s[0], s[1], k, True))
    assert sims[1].graph_stats()["host_finishes"] == 0
    assert sims[1].last_solve()["converged"] == 1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_step_is_the_host_step_in_the_pair_layouts(precision):
    case = _terrain(precision)
    print(f"[graph] terrain {precision} host vs device", flush=True)
    sims = _run(case, ("host", "device"), 5, every=lambda k, s: _same(s[0], s[1], k))
    assert sims[1].graph_stats()["host_finishes"] == 0
    assert sims[1].last_solve()["converged"] == 1


def test_graph_is_captured_again_after_a_configuration_change():
    from frehg2_amd.gw import RichardsSolver
    case = _columns("fp64")
    a, b = RichardsSolver(case), RichardsSolver(case)
    a.set_step_mode("device")
    b.set_step_mode("graph")
    for k in range(3):
        dt = a.next_dt()
        a.step(dt)
        b.step(dt)
    caps = b.graph_stats()["captures"]
    sink = np.zeros(a.field_shape(abi.F_HEAD))
    sink[1:-1, 1:-1, -20:] = 1e-7
    for s in (a, b):
        s.set_sink(sink)
    for k in range(3):
        dt = a.next_dt()
        a.step(dt)
        b.step(dt)
        _same(a, b, k)
    assert b.graph_stats()["captures"] == caps + 2
    assert a.audit()["tr_applied"] > 0.0
