"""Closed-form pins of the branches whole-step parity leaves open, run on the
oracle (CPU) and on the HIP path (gpu marker).

* Terrain-following gravity (Predictor.cpp:556-583, Corrector.cpp:194-250,
  :335-343): a saturated hydrostatic domain on a tilted, non-monotone bed is
  at rest.  The along-layer term sgn K r^2 v sin (predictor) / sgn K r v sin
  (corrector) must cancel the head difference K r v cos dh / dx of the
  hydrostatic heads exactly: with uniform layers the cell centres differ by
  dz_c and sin = |dz_c| / dist, cos = dx / dist (TerrainMetric.cpp:355-379),
  so q = (A K v r / dist)(-dz_c + sgn |dz_c|) vanishes iff sgn = sign(dz_c).
  A flipped sign in any of the four branches leaves a flux of 2 A K sin; an
  r instead of r^2 in the predictor (uniform density ratio c != 1) leaves a
  head drift.  A mirrored domain gives mirrored fields, which pins the plus-
  and minus-face sign rules against each other.
* Saturated contact, head rule: the reference's fringe tests
  (tests/unit/test_gw.cpp:1050-1079) on the post-allocation sweep alone.
* Adaptive dt (AdaptiveStep.cpp:59-117): in a 1D column without storage the
  controller's dq = |q_in - q_out| dtg / dz of a cell is its water-content
  change, so the shrink and keep branches are decided by max |dtheta| of the
  non-top cells; the Courant cap is courant_max dz / max dK/dtheta over the
  non-top unsaturated cells, with dK/dtheta from a central difference of
  K(h(theta)) (closures pinned by the reference's hand values).
"""
import numpy as np
import pytest

from frehg2_amd import abi
from frehg2_amd.cases import GwCase
from helpers import TEST_SOIL, gw_base_case, interior

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def make(backend, case):
    if backend == "oracle":
        from oracle.oracle import OracleRichards
        return OracleRichards(case)
    from frehg2_amd.gw import RichardsSolver
    return RichardsSolver(case)


# --------------------------------------------------------------------------
# terrain-following gravity at rest
# --------------------------------------------------------------------------
KS = 1.0e-5
LOAM = dict(vg_alpha=3.6, vg_n=1.56, theta_s=0.43, theta_r=0.078)


def _bed(nx, ny, dx, dy, flip_x=False, flip_y=False):
    """A bed that rises and falls along both axes (both slope signs on every
    axis) with a tilt that breaks the mirror symmetry."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    bx = 0.4 * np.sin(2 * np.pi * x / (nx * dx)) + 0.03 * x
    by = 0.25 * np.cos(2 * np.pi * y / (0.75 * ny * dy)) - 0.02 * y
    if flip_x:
        bx = bx[::-1]
    if flip_y:
        by = by[::-1]
    return by[:, None] + bx[None, :]


def _terrain_case(layers, pc="consistent", nx=8, ny=6, nz=10, ks=None):
    c = GwCase(name=f"terrain_{layers}_{pc}", nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=0.2,
               follow_terrain=True, terrain_layers=layers, bath=_bed(nx, ny, 2.0, 2.0),
               ksx=KS, ksy=KS, ksz=KS, aev=0.0, specific_storage=1e-4,
               predictor_corrector=pc, dt_init=300.0, dt_min=300.0, dt_max=300.0, dt=300.0,
               ic_form="head", ic_const=1.0, rtol=1e-13, atol=1e-300, max_iterations=4000,
               **LOAM)
    if layers == "scaled" or ks is not None:
        k = np.full((ny, nx, nz), KS) if ks is None else ks
        if layers == "scaled":
            # Scaled layers end every column on the box bottom, so the bottom
            # cells' bot3d differ only by rounding and the strict comparison
            # that picks the gravity sign (Predictor.cpp:560) is decided by
            # rounding there.  Their lateral faces are sealed (Ks = 0) so that
            # only layers with a real slope enter the pin.
            k = k.copy()
            k[:, :, -1] = 0.0
        c.ksx, c.ksy = k, k.copy()
    return c


def _install_hydrostatic(sim, ratio):
    """h + ratio z_c = const over the domain, saturated everywhere."""
    zc = sim.get(abi.F_BOT3D) + 0.5 * sim.get(abi.F_DZ3D)
    h = ratio * (float(np.max(zc)) - zc) + 0.5
    sim.set(abi.F_HEAD, h)
    sim.set(abi.F_WC, np.full_like(h, LOAM["theta_s"]))
    return interior(sim.get(abi.F_HEAD)).copy()


def _max_fluxes(sim):
    ax, ay = sim.get(abi.F_AREA_X), sim.get(abi.F_AREA_Y)
    qx, qy, qz = sim.get(abi.F_QX), sim.get(abi.F_QY), sim.get(abi.F_QZF)
    return (float(np.max(np.abs(qx))) / (KS * float(np.max(ax))),
            float(np.max(np.abs(qy))) / (KS * float(np.max(ay))),
            float(np.max(np.abs(qz))) / (KS * sim.case.dx * sim.case.dy))


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("layers,pc,ratios", [
    ("uniform", "consistent", None),
    ("scaled", "consistent", None),
    ("uniform", "legacy", None),
    ("uniform", "consistent", (1.025, 0.8)),
    ("scaled", "consistent", (0.97, 1.15)),
])
def test_terrain_hydrostatic_domain_stays_at_rest(backend, layers, pc, ratios):
    sim = make(backend, _terrain_case(layers, pc))
    if ratios is not None:
        shape = sim.field_shape(abi.F_HEAD)
        sim.set_density(np.full(shape, ratios[0]), np.full(shape, ratios[1]))
    # the pinned term is not vacuous: the along-layer slopes reach 10 % and
    # change sign along both axes
    sx, sy = interior(sim.get(abi.F_SIN_X)), interior(sim.get(abi.F_SIN_Y))
    assert sx.max() > 0.1 and sy.max() > 0.05
    bot = sim.get(abi.F_BOT3D)
    dbx = np.diff(bot[1:-1, 1:-1, 0], axis=1)
    dby = np.diff(bot[1:-1, 1:-1, 0], axis=0)
    assert (dbx > 0).any() and (dbx < 0).any() and (dby > 0).any() and (dby < 0).any()
    h0 = _install_hydrostatic(sim, 1.0 if ratios is None else ratios[0])
    worst_q, worst_h = 0.0, 0.0
    for n in range(10):
        sim.step()
        worst_q = max(worst_q, *_max_fluxes(sim))
        worst_h = max(worst_h, float(np.max(np.abs(interior(sim.get(abi.F_HEAD)) - h0))))
    # a wrong sign leaves fluxes of 2 sin ~ 0.2 Ks A; rounding leaves ~1e-16
    assert worst_q <= 1e-13, worst_q
    assert worst_h <= 1e-12, worst_h
    wc = interior(sim.get(abi.F_WC))
    assert np.max(np.abs(wc - LOAM["theta_s"])) <= 1e-14


@pytest.mark.parametrize("backend", BACKENDS)
def test_terrain_flipped_gravity_sign_is_not_at_rest(backend):
    """Anti-vacuity: the same domain with the hydrostatic heads of the mirrored
    bed (the sign of every slope reversed relative to the heads) moves."""
    sim = make(backend, _terrain_case("uniform"))
    zc = sim.get(abi.F_BOT3D) + 0.5 * sim.get(abi.F_DZ3D)
    zm = zc[:, ::-1, :]
    h = (float(np.max(zm)) - zm) + 0.5
    sim.set(abi.F_HEAD, h)
    sim.set(abi.F_WC, np.full_like(h, LOAM["theta_s"]))
    sim.step()
    assert max(_max_fluxes(sim)[:2]) > 1e-3


def _mirror_case(flip_x, flip_y, layers):
    nx, ny, nz = 9, 7, 12
    rng = np.random.Generator(np.random.Philox(20261203 + 77))
    ks = np.exp(np.log(5e-6) + 0.8 * rng.standard_normal((ny, nx, nz)))
    if flip_x:
        ks = ks[:, ::-1, :]
    if flip_y:
        ks = ks[::-1, :, :]
    c = _terrain_case(layers, nx=nx, ny=ny, nz=nz, ks=np.ascontiguousarray(ks))
    c.bath = _bed(nx, ny, 2.0, 2.0, flip_x, flip_y)
    c.name = f"mirror_{int(flip_x)}{int(flip_y)}_{layers}"
    c.ic_form, c.ic_const = "water_table", -1.2
    c.set_bc("top", abi.BC_FLUX, -4.0e-6)
    c.rtol = 1e-14
    return c


def _mirrored(field, flip_x, flip_y, face_x=False, face_y=False):
    """Map a halo-layout field of the mirrored run onto the base run's cells.
    Cell fields reverse over the halo; x (y) face fields, stored at the minus
    cell of their face, reverse over faces 0..nx (0..ny) and change sign."""
    f = field
    if flip_x:
        if face_x:
            f = np.concatenate([-f[:, -2::-1, :], f[:, -1:, :]], axis=1)
        else:
            f = f[:, ::-1, :]
    if flip_y:
        if face_y:
            f = np.concatenate([-f[-2::-1, :, :], f[-1:, :, :]], axis=0)
        else:
            f = f[::-1, :, :]
    return f


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("flip_x,flip_y,layers", [(True, False, "uniform"),
                                                  (False, True, "uniform"),
                                                  (True, True, "scaled")])
def test_mirrored_terrain_domain_gives_mirrored_fields(backend, flip_x, flip_y, layers):
    a = make(backend, _mirror_case(False, False, layers))
    b = make(backend, _mirror_case(flip_x, flip_y, layers))
    for n in range(6):
        a.step()
        b.step()
    assert a.audit()["boundary_in"] != 0.0
    checks = [(abi.F_HEAD, False, False), (abi.F_WC, False, False),
              (abi.F_QX, True, False), (abi.F_QY, False, True), (abi.F_QZF, False, False)]
    for fid, fx, fy in checks:
        fa = a.get(fid)[1:-1, 1:-1]
        fb = _mirrored(b.get(fid), flip_x, flip_y, fx, fy)[1:-1, 1:-1]
        den = float(np.max(np.abs(fa)))
        assert den > 0.0, fid
        err = float(np.max(np.abs(fa - fb))) / den
        assert err <= 1e-11, (fid, err)
    assert abs(a.volume() - b.volume()) <= 1e-13 * a.volume()


# --------------------------------------------------------------------------
# saturated contact: the reference's sc0 fringe (test_gw.cpp:990-1079)
# --------------------------------------------------------------------------
THETA_S = TEST_SOIL["theta_s"]


def _sweep_fringe_row(backend, axis, seed_at_minus):
    n = 6
    c = gw_base_case(n if axis == 0 else 1, n if axis == 1 else 1, n if axis == 2 else 1,
                     0.05, 0.5)
    sim = make(backend, c)
    sim.step()  # makes the face conductivities live
    shape = sim.field_shape(abi.F_HEAD)
    wc = np.full(shape, 0.1)
    h = np.full(shape, -10.0)

    def at(m):
        return (m + 1 if axis == 1 else 1, m + 1 if axis == 0 else 1, m if axis == 2 else 0)

    for m in range(n):
        wc[at(m)] = THETA_S - 1.0e-3
        h[at(m)] = 0.05
    wc[at(0 if seed_at_minus else n - 1)] = THETA_S
    sim.set(abi.F_HEAD, h)
    sim.set(abi.F_WC, wc)
    sim.test_post_allocate(0.5)
    wc_out, h_out = sim.get(abi.F_WC), sim.get(abi.F_HEAD)
    return [wc_out[at(m)] for m in range(n)], [h_out[at(m)] for m in range(n)]


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_head_classification_is_mirror_symmetric(backend, axis):
    # test_gw.cpp:1050-1063
    base_wc, base_h = _sweep_fringe_row(backend, axis, False)
    mir_wc, mir_h = _sweep_fringe_row(backend, axis, True)
    n = len(base_wc)
    for m in range(n):
        assert mir_wc[m] == base_wc[n - 1 - m], (axis, m)
        assert mir_h[m] == base_h[n - 1 - m], (axis, m)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("seed_at_minus", [False, True])
def test_head_classification_restores_the_whole_fringe(backend, axis, seed_at_minus):
    # test_gw.cpp:1065-1079
    wc, _ = _sweep_fringe_row(backend, axis, seed_at_minus)
    assert wc == [THETA_S] * len(wc), (axis, seed_at_minus, wc)


def _snapshot_pair(backend, axis, a_first):
    """Two live cells among dry ones: A (theta < theta_s, h above air entry:
    saturated by its head, with no saturated neighbour, so the sweep moves its
    head onto the retention curve) and B (unsaturated, below air entry)."""
    n = 4
    c = gw_base_case(n if axis == 0 else 1, n if axis == 1 else 1, n if axis == 2 else 1,
                     0.05, 0.5)
    sim = make(backend, c)
    sim.step()
    shape = sim.field_shape(abi.F_HEAD)
    wc = np.full(shape, 0.1)
    h = np.full(shape, -10.0)

    def at(m):
        return (m + 1 if axis == 1 else 1, m + 1 if axis == 0 else 1, m if axis == 2 else 0)

    pa, pb = (at(1), at(2)) if a_first else (at(2), at(1))
    wc[pa], h[pa] = 0.30, 0.05
    wc[pb], h[pb] = 0.25, -1.0
    sim.set(abi.F_HEAD, h)
    sim.set(abi.F_WC, wc)
    sim.test_post_allocate(0.5)
    w, hh = sim.get(abi.F_WC), sim.get(abi.F_HEAD)
    return (w[pa], hh[pa]), (w[pb], hh[pb])


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_head_rule_reads_the_pre_sweep_heads(backend, axis):
    """The head rule classifies against the state before the sweep
    (Reallocate.cpp:303-337 reads the snapshot): B sees A saturated by its
    pre-sweep head whichever of the two the sweep visits first, and takes
    theta(h_B); A, with no saturated neighbour, moves onto h(theta_A).  A sweep
    that read heads in place would leave B alone when A comes first."""
    from oracle import oracle as orc
    soil = (TEST_SOIL["vg_alpha"], TEST_SOIL["vg_n"], TEST_SOIL["theta_s"],
            TEST_SOIL["theta_r"], TEST_SOIL["aev"])
    for a_first in (True, False):
        (wa, ha), (wb, hb) = _snapshot_pair(backend, axis, a_first)
        assert hb == -1.0 and wa == 0.30
        assert wb == pytest.approx(orc.vg("theta", soil, -1.0), rel=1e-13), (axis, a_first)
        assert ha == pytest.approx(orc.vg("head", soil, 0.30), rel=1e-12), (axis, a_first)


def _oversaturated_state(sim):
    """A state the corrector could leave: an over-saturated block above a
    fringe just below theta_s, heads from the last step."""
    wc = sim.get(abi.F_WC)
    wc[1:-1, 1:-1, :4] = THETA_S + 2e-3
    wc[1:-1, 2:-1, 4:6] = THETA_S - 1e-4
    h = sim.get(abi.F_HEAD)
    h[1:-1, 1:-1, :3] = 0.02
    return wc, h


def _sweep_case():
    c = gw_base_case(4, 3, 12, 0.05, 0.5, ic_const=0.3)
    c.set_bc("top", abi.BC_FLUX, -2.0e-5)
    return c


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("surplus", ["drop", "redistribute"])
def test_post_allocation_sweep_closes_its_audit(backend, surplus):
    """The sweep alone moves no water through the boundary: the owned volume
    changes by realloc_adjust - vloss (RichardsSolver.hpp:86-130)."""
    c = _sweep_case()
    c.reallocation_surplus = surplus
    sim = make(backend, c)
    for n in range(3):
        sim.step()
    wc, h = _oversaturated_state(sim)
    sim.set(abi.F_HEAD, h)
    sim.set(abi.F_WC, wc)
    v0 = sim.volume()
    sim.test_post_allocate(0.5)
    a = sim.audit()
    v1 = sim.volume()
    if surplus == "drop":
        assert a["realloc_adjust"] < 0.0 and a["realloc_dropped"] > 0.0
    assert not np.array_equal(sim.get(abi.F_WC), wc)
    assert abs((v1 - v0) - (a["realloc_adjust"] - a["vloss"])) <= 1e-12 * v0
    assert interior(sim.get(abi.F_WC)).max() <= THETA_S


@pytest.mark.gpu
@pytest.mark.parametrize("surplus", ["drop", "redistribute"])
def test_gpu_post_allocation_sweep_matches_the_oracle(surplus):
    """The device sweep alone against the oracle's on one installed state."""
    from frehg2_amd.gw import RichardsSolver
    from oracle.oracle import OracleRichards
    c = _sweep_case()
    c.reallocation_surplus = surplus
    o, g = OracleRichards(c), RichardsSolver(c)
    for n in range(3):
        o.step()
        g.step()
    wc, h = _oversaturated_state(o)
    for sim in (o, g):
        sim.set(abi.F_HEAD, h)
        sim.set(abi.F_WC, wc)
        # identical face conductivities and fluxes going into the sweep
    for fid in (abi.F_KX, abi.F_KY, abi.F_KZF):
        assert np.max(np.abs(o.get(fid) - g.get(fid))) <= 1e-10 * np.max(np.abs(o.get(fid)))
    o.test_post_allocate(0.5)
    g.test_post_allocate(0.5)
    for fid in (abi.F_WC, abi.F_HEAD, abi.F_QZF):
        a, b = g.get(fid), o.get(fid)
        assert np.max(np.abs(a - b)) <= 1e-12 * max(np.max(np.abs(b)), 1e-30), fid
    # audit terms relative to the over-saturated volume the sweep moves
    w = interior(wc)
    moved = float(np.sum(np.maximum(w - THETA_S, 0.0))) * c.dx * c.dy * c.dz
    assert moved > 0.0
    ao, ag = o.audit(), g.audit()
    for k in ("realloc_adjust", "realloc_dropped", "vloss"):
        assert abs(ao[k] - ag[k]) <= 1e-12 * moved, (k, ao[k], ag[k])


# --------------------------------------------------------------------------
# adaptive dt: shrink / keep / grow and the Courant cap (AdaptiveStep.cpp)
# --------------------------------------------------------------------------
def _column(**kw):
    opts = dict(dt_min=1e-3, dt_max=1e6, dq_grow=1e30, dq_shrink=2e30, courant_max=1e30)
    opts.update(kw)
    c = gw_base_case(1, 1, 20, 0.05, 60.0, **opts)
    c.set_bc("top", abi.BC_FLUX, -3.0e-6)
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)
    c.adaptive = True
    return c


def _one_step_dtheta(backend, **kw):
    sim = make(backend, _column(**kw))
    wc0 = interior(sim.get(abi.F_WC))[0, 0].copy()
    dtg = sim.step()
    dwc = np.abs(interior(sim.get(abi.F_WC))[0, 0] - wc0)
    return sim, dtg, dwc


@pytest.mark.parametrize("backend", BACKENDS)
def test_adaptive_dt_shrinks_on_the_non_top_moisture_change(backend):
    probe, dtg, dwc = _one_step_dtheta(backend)
    d = float(dwc[1:].max())
    assert d > 0.0
    # the top cell changes most: counting it would decide otherwise
    assert dwc[0] > 2.0 * d
    cases = [  # (dq_grow, dq_shrink) -> factor
        (0.5 * d, 0.999 * d, 0.75),     # dq > dq_shrink: shrink
        (0.5 * d, 1.001 * d, 1.0),      # between: keep
        (1.001 * d, 2.0 * d, 1.25),     # dq < dq_grow: grow
        (0.5 * d, 1.001 * dwc[0], 1.0)  # the top cell's change is not counted
    ]
    for grow, shrink, factor in cases:
        sim, dt2, _ = _one_step_dtheta(backend, dq_grow=grow, dq_shrink=shrink)
        assert dt2 == dtg
        assert sim.next_dt() == pytest.approx(factor * dtg, rel=1e-15), (grow, shrink, factor)


@pytest.mark.parametrize("backend", BACKENDS)
def test_adaptive_dt_courant_cap(backend):
    from oracle import oracle as orc
    probe, dtg, _ = _one_step_dtheta(backend)
    wc = interior(probe.get(abi.F_WC))[0, 0]
    soil = (TEST_SOIL["vg_alpha"], TEST_SOIL["vg_n"], TEST_SOIL["theta_s"],
            TEST_SOIL["theta_r"], TEST_SOIL["aev"])
    ks = probe.case.ksz
    dz = probe.case.dz

    def k_of_theta(w):
        return orc.vg("conductivity", soil, orc.vg("head", soil, w), ks)

    dk = []
    for k in range(1, len(wc)):  # non-top cells, all unsaturated here
        assert wc[k] < 0.99 * soil[2]
        e = 1e-6 * (soil[2] - soil[3])
        dk.append((k_of_theta(wc[k] + e) - k_of_theta(wc[k] - e)) / (2 * e))
    cap1 = dz / max(dk)  # Courant dt at courant_max = 1
    co = 0.5 * 1.25 * dtg / cap1
    sim, dt2, _ = _one_step_dtheta(backend, courant_max=co)
    assert dt2 == dtg
    assert sim.next_dt() == pytest.approx(co * cap1, rel=1e-6)
    assert sim.next_dt() < dtg
    # without the cap the same step grows
    assert probe.next_dt() == pytest.approx(1.25 * dtg, rel=1e-15)
