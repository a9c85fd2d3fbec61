"""Batched direct column solve (laterally decoupled operators).

With no lateral conductivity (the 4096-column workload: ksx = ksy = 0) the
operator is a batch of independent tridiagonal columns, which the reference
solves like any other system with KSPCG (src/core/LinearSystem.cpp:320).
The z-line form here solves every column directly in one launch: a wave a
column over its active rows, the pivots by a continuant scan over the lanes,
the LDL^T recurrences, one FP64 refinement, and a residual check against the
PCG tolerance (the PCG finishes a solve that misses it).

Pins:
* the step with the direct solve against the same step with the z-line PCG
  (frehg_gw_enable_column_solve(0)) from identical inputs, FP64: head and
  theta within 1e-11 normwise (both solves stop at a residual of
  1e-12 ||b||, and theta amplifies head differences near saturation); the
  direct step against the oracle's step within 1e-10 in both forms (the
  column solve is FP64 in the mixed form too);
* a deep saturated hydrostatic column at rest stays at rest (the stiffest
  operator of the workload, the hard case for the scan's pivots), with the
  direct solve and with the PCG;
* the tile of each depth bucket changes which wave solves which column, not
  the arithmetic: the solution is bitwise the same for every tile;
* the depth buckets: R = ceil(active rows / 64), deepest first, the counts
  summing to the columns with active rows;
* tune_columns leaves a valid tile per bucket and the solution unchanged.
"""
import numpy as np
import pytest

from frehg2_amd import abi, cases
from helpers import interior

pytestmark = pytest.mark.gpu


def _columns(side=16, lmax=400, precision="fp64"):
    c = cases.columns_4096(ncol_side=side, lmax=lmax)
    c.precision = precision
    c.rtol, c.atol = 1e-12, 1e-300
    return c


def _nerr(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def _steps(sim, case, n):
    for k in range(n):
        for which, vals in case.bc_schedule(k).items():
            sim.set_bc_values(which, vals)
        sim.step(sim.next_dt())


def test_buckets_are_the_column_depths():
    from frehg2_amd.gw import RichardsSolver
    c = _columns()
    g = RichardsSolver(c)
    tiles = g.column_tiles()
    depth = c.depth  # active rows of each column
    want = {}
    for L in depth.ravel():
        R = (int(L) + 63) // 64
        want[R] = want.get(R, 0) + 1
    assert [t["R"] for t in tiles] == sorted(want, reverse=True)
    assert {t["R"]: t["columns"] for t in tiles} == want
    assert all(t["cpw"] == 1 for t in tiles)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_direct_solve_is_the_pcg_solve(precision):
    from frehg2_amd.gw import RichardsSolver
    from oracle.oracle import OracleRichards
    c = _columns(precision=precision)
    d = RichardsSolver(c)
    assert d.column_tiles()
    p = RichardsSolver(c)
    p.enable_column_solve(False)
    o = OracleRichards(c)
    for k in range(6):
        for which, vals in c.bc_schedule(k).items():
            for s in (d, p, o):
                s.set_bc_values(which, vals)
        for s in (d, p):
            s.set(abi.F_HEAD, o.get(abi.F_HEAD))
            s.set(abi.F_WC, o.get(abi.F_WC))
        dt = o.next_dt()
        for s in (d, p, o):
            s.step(dt)
        st = d.last_solve()
        assert st["converged"] == 1 and st["iterations"] == 1, st
        assert p.last_solve()["converged"] == 1
        for f in (abi.F_HEAD, abi.F_WC):
            a, b, r = interior(d.get(f)), interior(p.get(f)), interior(o.get(f))
            assert _nerr(a, r) <= 1e-10, (k, f, _nerr(a, r))
            if precision == "fp64":
                assert _nerr(a, b) <= 1e-11, (k, f, _nerr(a, b))


def test_tiles_do_not_change_the_solution():
    from frehg2_amd.gw import RichardsSolver
    c = _columns()
    runs = []
    for tiles in (None, "wide"):
        g = RichardsSolver(c)
        if tiles == "wide":
            g.set_column_tiles([min(8, t["columns"]) for t in g.column_tiles()])
            assert all(t["cpw"] > 1 or t["columns"] == 1 for t in g.column_tiles())
        _steps(g, c, 4)
        runs.append((g.get(abi.F_HEAD), g.get(abi.F_WC)))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    with pytest.raises(abi.FrehgError):
        g.set_column_tiles([1])  # one tile per bucket
    with pytest.raises(abi.FrehgError):
        g.set_column_tiles([0] * len(g.column_tiles()))


def test_tuned_tiles_keep_the_solution():
    from frehg2_amd import autotune
    from frehg2_amd.gw import RichardsSolver
    c = _columns()
    a, b = RichardsSolver(c), RichardsSolver(c)
    _steps(a, c, 2)
    _steps(b, c, 2)
    res = autotune.tune_columns(b, repeats=1, candidates=(1, 4))
    print(f"[columns] tuned tiles {res['tiles']}", flush=True)
    assert [t["cpw"] for t in b.column_tiles()] == [t["cpw"] for t in res["tiles"]]
    for k in range(2, 4):
        for which, vals in c.bc_schedule(k).items():
            a.set_bc_values(which, vals)
            b.set_bc_values(which, vals)
        dt = a.next_dt()  # b's controller holds the trial step's value
        a.step(dt)
        b.step(dt)
    assert np.array_equal(a.get(abi.F_HEAD), b.get(abi.F_HEAD))
    assert np.array_equal(a.get(abi.F_VLOSS), b.get(abi.F_VLOSS))


@pytest.mark.parametrize("direct", [True, False])
def test_deep_saturated_column_stays_at_rest(direct):
    """A 1x1x400 column saturated and hydrostatic (water table 5 m above its
    top), no flux in or out: h + z is constant, no water moves, and the head
    system's solution is the state itself.  Its operator is the stiff one
    (storage Ss dz / dt against K / dz: a condition near 1e9), the hard case
    for the scan's pivots; the direct solve and the PCG both keep it at
    rest."""
    from frehg2_amd.gw import RichardsSolver
    c = cases.columns_4096(ncol_side=1, lmax=400)
    c.depth = np.array([[400]])
    c.bath = np.zeros((1, 1))
    c.ic_form, c.ic_values, c.ic_const = "water_table", None, 5.0
    c.set_bc("top", abi.BC_FLUX, 0.0)
    c.set_bc("bot", abi.BC_NOFLUX, 0.0)
    c.bc_schedule = None
    c.adaptive = False
    # the workload's tolerance: at 1e-12 the residual of this operator sits
    # at its rounding floor (||A|| ||x|| eps over 400 rows) for any solver
    c.rtol, c.atol = 1e-10, 1e-300
    g = RichardsSolver(c)
    assert [t["R"] for t in g.column_tiles()] == [7]
    g.enable_column_solve(direct)
    h0 = g.get(abi.F_HEAD)
    assert float(np.min(interior(h0))) > 0.0
    for _ in range(3):
        g.step(60.0)
        assert g.last_solve()["converged"] == 1
    err = _nerr(interior(g.get(abi.F_HEAD)), interior(h0))
    print(f"[columns] at rest, direct {direct}: {err:.2e}", flush=True)
    assert err <= 1e-10
