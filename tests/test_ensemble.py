"""Ensembles stacked along y with inactive separator rows.

The claim under test: a row of fully inactive columns between members is an
exact no-flux wall, so a stacked run advances every member as it would run
alone at the same dt.  CPU tests check the layout, the construction checks
and the claim on the oracle; GPU tests check it on the HIP path (FP64 and
mixed) and hold the stacked GPU step to the oracle's on identical inputs.
"""
import numpy as np
import pytest

from frehg2_amd import abi, cases
from frehg2_amd.ensemble import Ensemble
from helpers import interior

CPL = ("eta", "dept", "avail", "seep_accum", "gain")


def _members(count=3, nx=10, ny=7, nz=16, precision="fp64", rtol=1e-13):
    out = []
    for m in range(count):
        c = cases.coupled_128x128x64(member=m, nx=nx, ny=ny, nz=nz, precision=precision)
        c.rtol, c.atol, c.max_iterations = rtol, 1e-300, 5000
        out.append(c)
    return out


def _nerr(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def _member_errors(ens, stacked, solos):
    """Worst normwise error over members of h, theta and coupling state."""
    h, w = stacked.get(abi.F_HEAD), stacked.get(abi.F_WC)
    worst = {"head": 0.0, "theta": 0.0, "coupling": 0.0}
    for m, s in enumerate(solos):
        worst["head"] = max(worst["head"], _nerr(ens.member_view(h, m), interior(s.get(abi.F_HEAD))))
        worst["theta"] = max(worst["theta"], _nerr(ens.member_view(w, m), interior(s.get(abi.F_WC))))
        for n in CPL:
            a = ens.member_coupling(stacked.coupling_state(n), m)[1:-1, 1:-1]
            b = s.coupling_state(n)[1:-1, 1:-1]
            worst["coupling"] = max(worst["coupling"],
                                    float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-12))
    return worst


def _run_pair(make, ens, solos, steps):
    stacked = make(ens.case)
    runs = [make(c) for c in solos]
    worst = {}
    vol_err = 0.0
    for s in range(steps):
        dt = stacked.next_dt()
        stacked.step(dt)
        for r in runs:
            r.step(dt)
        e = _member_errors(ens, stacked, runs)
        worst = {k: max(worst.get(k, 0.0), v) for k, v in e.items()}
        vs, vm = stacked.volume(), sum(r.volume() for r in runs)
        vol_err = max(vol_err, abs(vs - vm) / abs(vm))
    return worst, vol_err


# ---- CPU --------------------------------------------------------------------

def test_layout_separators_are_fully_inactive():
    mem = _members()
    ens = Ensemble(mem)
    c = ens.case
    assert c.ny == 3 * 7 + 2
    top = max(float(np.max(m.bath if m.bath is not None else m.bath_const)) for m in mem)
    depth = c.dz * c.nz
    for r in (7, 15):
        assert np.all(c.bath[r] < top - depth)         # the bed lies below the mesh bottom
    for m in range(3):
        np.testing.assert_array_equal(c.ksz[ens.rows(m)], mem[m].soil_array("ksz"))
        cp = ens.member_coupling(c.coupling["dept"], m)
        np.testing.assert_array_equal(cp[1:-1, 1:-1], mem[m].coupling["dept"][1:-1, 1:-1])
    h = np.zeros((c.ny + 2, c.nx + 2, c.nz))
    h[1 + ens.rows(1).start, 1, 0] = 7.0
    assert ens.member_view(h, 1)[0, 0, 0] == 7.0


def test_incompatible_members_are_rejected():
    a, b = _members(2)
    with pytest.raises(ValueError, match="nz"):
        Ensemble([a, cases.coupled_128x128x64(member=1, nx=10, ny=7, nz=12)])
    b.set_bc("ym", abi.BC_HEAD, 0.0)
    with pytest.raises(ValueError, match="no-flux y"):
        Ensemble([a, b])
    c = cases.loam_32x32x50(nx=10, ny=7, nz=16)
    d = cases.loam_32x32x50(nx=10, ny=7, nz=16)
    d.bath_const = 0.5
    with pytest.raises(ValueError, match="highest bed"):
        Ensemble([c, d])
    e = cases.hetero_512x512x100(nx=10, ny=7, nz=16)
    with pytest.raises(ValueError, match="regular"):
        Ensemble([e, e])
    f = _members(1)[0]
    f.coupling = None
    with pytest.raises(ValueError, match="coupled"):
        Ensemble([a, f])


def test_stacked_schedule_places_member_values():
    mem = []
    for m in range(2):
        c = cases.loam_32x32x50(nx=4, ny=3, nz=6)
        c.set_bc("top", abi.BC_FLUX, -1e-6)
        c.bc_schedule = (lambda v: (lambda step: {0: np.full((3, 4), v * (step + 1))}))(m + 1.0)
        mem.append(c)
    ens = Ensemble(mem)
    got = ens.case.bc_schedule(2)[0]
    assert got.shape == (7, 4)
    np.testing.assert_array_equal(got[0:3], 3.0)
    np.testing.assert_array_equal(got[4:7], 6.0)
    assert np.all(got[3] == 0.0)                       # separator row: no-flux, value unused


def test_oracle_stacked_run_equals_member_runs():
    from oracle.oracle import OracleRichards
    mem = _members()
    worst, vol_err = _run_pair(OracleRichards, Ensemble(mem), mem, 4)
    assert worst["head"] <= 1e-10 and worst["theta"] <= 1e-10, worst
    assert worst["coupling"] <= 1e-10, worst
    assert vol_err <= 1e-13


# ---- GPU --------------------------------------------------------------------

@pytest.mark.gpu
def test_gpu_stacked_run_equals_member_runs():
    from frehg2_amd.gw import RichardsSolver
    mem = _members(count=4, nx=16, ny=12, nz=20)
    print("[ensemble] fp64 4 x (16,12,20)", flush=True)
    worst, vol_err = _run_pair(RichardsSolver, Ensemble(mem), mem, 6)
    assert worst["head"] <= 1e-10 and worst["theta"] <= 1e-10, worst
    assert worst["coupling"] <= 1e-10, worst
    assert vol_err <= 1e-13


@pytest.mark.gpu
def test_gpu_mixed_stacked_run_against_fp64_members():
    from frehg2_amd.gw import RichardsSolver
    mem = _members(count=4, nx=16, ny=12, nz=20, rtol=1e-12)
    mixed = _members(count=4, nx=16, ny=12, nz=20, precision="mixed", rtol=1e-12)
    ens = Ensemble(mixed)
    print("[ensemble] mixed 4 x (16,12,20) vs fp64 members", flush=True)
    stacked = RichardsSolver(ens.case)
    runs = [RichardsSolver(c) for c in mem]
    for s in range(6):
        dt = stacked.next_dt()
        stacked.step(dt)
        for r in runs:
            r.step(dt)
        e = _member_errors(ens, stacked, runs)
        assert e["head"] <= 1e-8 and e["theta"] <= 1e-8, (s, e)
        assert e["coupling"] <= 1e-8, (s, e)


@pytest.mark.gpu
def test_gpu_stacked_step_matches_oracle():
    """Identical inputs each step: the stacked GPU step against the stacked
    oracle step, 1e-10 normwise."""
    from frehg2_amd.gw import RichardsSolver
    from oracle.oracle import OracleRichards
    case = Ensemble(_members(count=3, nx=12, ny=9, nz=18, rtol=1e-12)).case
    print(f"[ensemble] stacked parity ({case.nx},{case.ny},{case.nz})", flush=True)
    o, g = OracleRichards(case), RichardsSolver(case)
    for s in range(5):
        g.set(abi.F_HEAD, o.get(abi.F_HEAD))
        g.set(abi.F_WC, o.get(abi.F_WC))
        for n in CPL:
            g.set_coupling_state(n, o.coupling_state(n))
        dt = o.next_dt()
        o.step(dt)
        g.step(dt)
        for f in (abi.F_HEAD, abi.F_WC):
            assert _nerr(interior(g.get(f)), interior(o.get(f))) <= 1e-10, (s, f)
        for n in CPL:
            a, b = g.coupling_state(n), o.coupling_state(n)
            assert np.max(np.abs(a - b)) <= 1e-10 * max(np.max(np.abs(b)), 1e-12), (s, n)
        assert abs(g.volume() - o.volume()) <= 1e-12 * abs(o.volume())
