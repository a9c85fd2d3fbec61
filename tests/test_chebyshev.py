"""Inner preconditioners of the mixed form: point Jacobi and the FP32
Jacobi-Chebyshev polynomial (FREHG_PC_CHEBYSHEV, frehg_gw_set_chebyshev)
beside the z-line default.

The reference's KSPCG takes point Jacobi, block Jacobi / ICC or AMG
(src/core/LinearSystem.cpp:158-221); any of them changes the iteration, not
the converged solution.  So every form is held to the FP64 oracle's step
(zline; the oracle has no polynomial preconditioner) from identical inputs at
the mixed form's bound, head and theta within 1e-8 normwise, on the stiff
32x32x50 workload and on a terrain-following slice of the 26.2 M-cell one,
with the audit closing; the Chebyshev kernels run inside a captured step
bitwise as outside; the options are validated.
"""
import dataclasses

import numpy as np
import pytest

from frehg2_amd import abi, cases
from helpers import interior

WORKLOADS = {
    "stiff_32x32x50": lambda: (cases.stiff_32x32x50(precision="mixed"), 1.0),
    "hetero_slice_32x24x100": lambda: (cases.hetero_512x512x100(nx=32, ny=24, nz=100,
                                                                precision="mixed"), 300.0),
}


def _nerr(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def _variant(case, pc, precision="mixed"):
    c = dataclasses.replace(case)
    c.preconditioner = pc
    c.precision = precision
    c.rtol, c.atol = 1e-11, 1e-300
    return c


def test_options_are_validated():
    from frehg2_amd.gw import RichardsSolver
    case, _ = WORKLOADS["stiff_32x32x50"]()
    with pytest.raises(abi.FrehgError) as e:  # the polynomial is the FP32 form's
        RichardsSolver(_variant(case, "chebyshev", precision="fp64"))
    assert e.value.code == abi.FREHG_EINVAL and "Chebyshev" in str(e.value)
    with pytest.raises(abi.FrehgError):  # the oracle has no polynomial preconditioner
        from oracle.oracle import OracleRichards
        OracleRichards(_variant(case, "chebyshev"))


@pytest.mark.gpu
def test_chebyshev_options_on_the_device():
    from frehg2_amd.gw import RichardsSolver
    case, _ = WORKLOADS["stiff_32x32x50"]()
    g = RichardsSolver(_variant(case, "chebyshev"))
    for bad in ((4, 0.05), (1, 0.05), (3, 0.0), (3, 1.0)):
        with pytest.raises(abi.FrehgError):
            g.set_chebyshev(*bad)
    g.set_chebyshev(5, 0.1)
    z = RichardsSolver(_variant(case, "zline"))
    with pytest.raises(abi.FrehgError):
        z.set_chebyshev(3, 0.05)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(WORKLOADS))
@pytest.mark.parametrize("pc,degree", [("jacobi", 0), ("chebyshev", 3), ("chebyshev", 5),
                                       ("zline", 0)])
def test_mixed_inner_preconditioners_meet_the_oracle(name, pc, degree):
    from frehg2_amd.gw import RichardsSolver
    from oracle.oracle import OracleRichards
    case, dt = WORKLOADS[name]()
    g = RichardsSolver(_variant(case, pc))
    if pc == "chebyshev":
        g.set_chebyshev(degree, 0.05)
    o = OracleRichards(_variant(case, "zline", precision="fp64"))
    worst = 0.0
    its = 0
    for k in range(3):
        g.set(abi.F_HEAD, o.get(abi.F_HEAD))
        g.set(abi.F_WC, o.get(abi.F_WC))
        step_dt = o.next_dt() if case.adaptive else dt
        v0 = g.volume()
        o.step(step_dt)
        g.step(step_dt)
        st = g.last_solve()
        assert st["converged"] == 1, st
        its += st["iterations"]
        for f in (abi.F_HEAD, abi.F_WC):
            e = _nerr(interior(g.get(f)), interior(o.get(f)))
            worst = max(worst, e)
            assert e <= 1e-8, (name, pc, degree, k, f, e)
        a = g.audit()
        res = (g.volume() - v0) - (a["boundary_in"] - a["ss_storage"] + a["realloc_adjust"]
                                   - a["vloss"])
        assert abs(res) <= 1e-12 * g.volume(), res
    print(f"[pc] {name} {pc}{degree or ''}: {its / 3:.1f} it/solve, worst {worst:.1e}", flush=True)


@pytest.mark.gpu
def test_chebyshev_inside_a_captured_step():
    from frehg2_amd.gw import RichardsSolver
    case, _ = WORKLOADS["stiff_32x32x50"]()
    c = _variant(case, "chebyshev")
    a, b = RichardsSolver(c), RichardsSolver(c)
    a.set_step_mode("device")
    b.set_step_mode("graph")
    for k in range(5):
        dt = a.next_dt()
        a.step(dt)
        b.step(dt)
        for f in (abi.F_HEAD, abi.F_WC):
            assert np.array_equal(a.get(f), b.get(f)), (k, f)
        assert a.last_solve() == b.last_solve()
    assert b.graph_stats()["replays"] == 4
