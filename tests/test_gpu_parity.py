"""Parity of the HIP step with the FP64 oracle on the same seeded inputs.

Bar (normwise, as pointwise relative error is undefined where h ~ 0):
||h_gpu - h_oracle||_inf / ||h_oracle||_inf <= 1e-10, the same for theta,
audit terms within 1e-10 of the larger magnitude, every step.  The mixed
form is held to 1e-8 against the FP64 oracle.  Solver tolerances are
tightened so the comparison measures the step, not two solver stopping
points.
"""
import numpy as np
import pytest

from frehg2_amd import abi, cases
from helpers import gw_base_case, interior

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _pair(case):
    from frehg2_amd.gw import RichardsSolver
    from oracle.oracle import OracleRichards
    print(f"[parity] case={case.name} shape=({case.nx},{case.ny},{case.nz}) "
          f"pc={case.preconditioner} prec={case.precision}", flush=True)
    return OracleRichards(case), RichardsSolver(case)


def _nerr(a, b):
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den


def _compare(o, g, tag, tol=TOL, audit_tol=1e-10):
    eh = _nerr(interior(g.get(abi.F_HEAD)), interior(o.get(abi.F_HEAD)))
    ew = _nerr(interior(g.get(abi.F_WC)), interior(o.get(abi.F_WC)))
    assert eh <= tol, (tag, "head", eh)
    assert ew <= tol, (tag, "theta", ew)
    ao, ag = o.audit(), g.audit()
    scale = max(abs(v) for v in ao.values()) + 1e-300
    for k in ao:
        assert abs(ao[k] - ag[k]) <= audit_tol * scale + 1e-18, (tag, k, ao[k], ag[k])
    vo, vg = o.volume(), g.volume()
    assert abs(vo - vg) <= 1e-12 * abs(vo), (tag, "volume", vo, vg)
    assert g.next_dt() == pytest.approx(o.next_dt(), rel=1e-12)
    return eh, ew


def _run(case, steps, tol=TOL, tighten=True, audit_tol=1e-10):
    if tighten:
        case.rtol, case.atol = 1e-13, 1e-300
        case.max_iterations = 5000
    o, g = _pair(case)
    worst = (0.0, 0.0)
    for s in range(steps):
        if case.bc_schedule is not None:
            for which, vals in case.bc_schedule(s).items():
                o.set_bc_values(which, vals)
                g.set_bc_values(which, vals)
        dt = case.dt
        o.step(dt)
        g.step(dt)
        eh, ew = _compare(o, g, f"{case.name} step {s}", tol, audit_tol)
        worst = (max(worst[0], eh), max(worst[1], ew))
    print(f"[parity] {case.name}: worst head {worst[0]:.3e} theta {worst[1]:.3e} "
          f"iters(gpu) {g.last_solve()['iterations']}", flush=True)
    return o, g


def test_initial_state_and_geometry_match():
    c = cases.hetero_512x512x100(nx=24, ny=20, nz=30)
    o, g = _pair(c)
    for f in (abi.F_HEAD, abi.F_WC, abi.F_DZ3D, abi.F_BOT3D, abi.F_AREA_X, abi.F_AREA_Y,
              abi.F_SIN_X, abi.F_COS_X, abi.F_SIN_Y, abi.F_COS_Y):
        a, b = g.get(f), o.get(f)
        assert _nerr(a, b) <= 1e-14, f
    assert g.active_cells() == o.active_cells()


def test_loam_trajectory():
    _run(cases.loam_32x32x50(), 20)


def test_loam_jacobi_preconditioner():
    c = cases.loam_32x32x50(nx=8, ny=6)
    c.preconditioner = "jacobi"
    _run(c, 10)


def test_stiff_fp64_trajectory():
    _run(cases.stiff_32x32x50(precision="fp64"), 15)


def test_hetero_terrain_reduced():
    _run(cases.hetero_512x512x100(nx=24, ny=20, nz=30, steps=5), 5)


def test_columns_reduced_with_rain_schedule():
    c = cases.columns_4096(ncol_side=8, lmax=120)
    o, g = _run(c, 12)
    assert o.active_cells() == int(c.depth.clip(max=120).sum()) or o.active_cells() > 0


def test_coupled_reduced():
    c = cases.coupled_128x128x64(nx=16, ny=12, nz=20, precision="fp64")
    o, g = _run(c, 8)
    for name in ("eta", "dept", "avail", "seep_accum", "gain"):
        a, b = g.coupling_state(name), o.coupling_state(name)
        assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1e-12), name


@pytest.mark.parametrize("variant", ["harmonic", "redistribute", "disabled", "full3d_off"])
def test_scheme_variants(variant):
    c = cases.stiff_32x32x50(precision="fp64")
    c.nx = c.ny = 10
    for n in ("ksx", "ksy", "ksz", "vg_alpha", "vg_n"):
        setattr(c, n, np.asarray(getattr(c, n))[:10, :10, :])
    if variant == "harmonic":
        c.face_conductivity = "harmonic"
    elif variant == "redistribute":
        c.reallocation_surplus = "redistribute"
    elif variant == "disabled":
        c.reallocation = "disabled"
    else:
        c.use_full3d = False
    _run(c, 8)


def test_side_boundaries_and_masking():
    # hydrostatic and flux sides, a head bottom and masked columns together
    bath = np.zeros((6, 7))
    bath[2, 3] = -1.0
    bath[0, 0] = -0.13
    c = gw_base_case(7, 6, 10, 0.1, 5.0, specific_storage=1e-5, ic_form="water_table",
                     ic_const=-0.4, bath=bath)
    c.set_bc("xm", abi.BC_HEAD_HYDROSTATIC, -0.3)
    c.set_bc("xp", abi.BC_FLUX, 2e-6)
    c.set_bc("ym", abi.BC_HEAD, -0.6)
    c.set_bc("top", abi.BC_FLUX, -3e-6)
    c.set_bc("bot", abi.BC_HEAD, 0.7)
    _run(c, 10)


def test_mixed_precision_against_fp64_oracle():
    c = cases.stiff_32x32x50(precision="mixed")
    c.rtol, c.atol = 1e-12, 1e-300
    o, g = _pair(c)
    worst = 0.0
    for s in range(10):
        o.step(c.dt)
        g.step(c.dt)
        eh = _nerr(interior(g.get(abi.F_HEAD)), interior(o.get(abi.F_HEAD)))
        ew = _nerr(interior(g.get(abi.F_WC)), interior(o.get(abi.F_WC)))
        worst = max(worst, eh, ew)
        assert eh <= 1e-8 and ew <= 1e-8, (s, eh, ew)
        st = g.last_solve()
        assert st["converged"] == 1 and st["refinements"] >= 1
    print(f"[parity] mixed worst normwise error {worst:.3e}", flush=True)


def test_audit_identity_on_gpu_columns():
    c = cases.columns_4096(ncol_side=8, lmax=200)
    from frehg2_amd.gw import RichardsSolver
    g = RichardsSolver(c)
    prev = g.volume()
    for s in range(10):
        for which, vals in c.bc_schedule(s).items():
            g.set_bc_values(which, vals)
        g.step(c.dt)
        v = g.volume()
        a = g.audit()
        res = (v - prev) - (a["boundary_in"] - a["ss_storage"] + a["realloc_adjust"] - a["vloss"])
        assert abs(res) <= 1e-12 * max(v, 1.0), (s, res)
        prev = v


# ---- the solve seam (LinearSystem) on the GPU -------------------------------
def _poisson(split):
    n = 101
    hh = 1.0 / (n + 1)
    inv = 1.0 / (hh * hh)
    rows, cols, vals = [], [], []
    for r in range(n):
        rows.append(r), cols.append(r), vals.append(2.0 * inv)
        w = r - 1 if r > 0 else -1
        e = r + 1 if r < n - 1 else -1
        legs = [(w, -0.5 * inv)] * 2 + [(e, -0.5 * inv)] * 2 if split else [(w, -inv), (e, -inv)]
        for cc, v in legs:
            rows.append(r), cols.append(cc), vals.append(v)
    b = np.zeros(n)
    b[0] += inv
    b[-1] += 2.0 * inv
    return n, np.array(rows), np.array(cols), np.array(vals), b, 1.0 + (np.arange(n) + 1) * hh


@pytest.mark.parametrize("split", [False, True])
def test_ls_seam_poisson(split):
    import torch
    from frehg2_amd.gw import LinearSystem
    n, rows, cols, vals, b, exact = _poisson(split)
    ls = LinearSystem(n, rtol=1e-12, atol=1e-30, max_iterations=5000)
    ls.set_pattern(rows, cols)
    vd = torch.tensor(vals, device="cuda")
    ls.set_values(vd)
    assert ls.is_symmetric(1e-12)
    bd = torch.tensor(b, device="cuda")
    xd = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = ls.solve(bd, xd)
    assert st["converged"] == 1 and st["iterations"] > 0
    assert np.max(np.abs(xd.cpu().numpy() - exact)) < 1e-9


def test_ls_seam_errors():
    import torch
    from frehg2_amd.gw import LinearSystem
    n, rows, cols, vals, b, _ = _poisson(False)
    vals = vals.copy()
    vals[2] *= 1.5
    ls = LinearSystem(n, rtol=1e-14, atol=1e-300, max_iterations=1)
    with pytest.raises(abi.FrehgError):
        ls.set_values(torch.tensor(vals, device="cuda"))  # before set_pattern
    ls.set_pattern(rows, cols)
    ls.set_values(torch.tensor(vals, device="cuda"))
    assert not ls.is_symmetric(1e-12)
    with pytest.raises(abi.FrehgError) as e:
        ls.solve(torch.tensor(b, device="cuda"), torch.zeros(n, dtype=torch.float64, device="cuda"))
    assert e.value.code == abi.FREHG_EDIVERGED
