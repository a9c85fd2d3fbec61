"""The reference's groundwater invariants (tests/unit/test_gw.cpp), run on the
oracle (CPU) and on the HIP path (gpu marker).

Each test cites the reference test it ports, with the configuration it names
(predictor_corrector: legacy where the reference asks for the legacy pair).
"""
import numpy as np
import pytest

from frehg2_amd import abi
from helpers import TEST_SOIL, gw_base_case, interior, run_steps

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def make(backend, case):
    if backend == "oracle":
        from oracle.oracle import OracleRichards
        return OracleRichards(case)
    from frehg2_amd.gw import RichardsSolver
    return RichardsSolver(case)


def closure(sim, prev):
    v = sim.volume()
    a = sim.audit()
    return v, (v - prev) - (a["boundary_in"] - a["ss_storage"] + a["realloc_adjust"] - a["vloss"])


@pytest.mark.parametrize("backend", BACKENDS)
def test_closed_column_conserves_mass(backend):
    # test_gw.cpp:147-166
    sim = make(backend, gw_base_case(1, 1, 20, 0.05, 0.5))
    v0 = prev = sim.volume()
    for n in range(50):
        sim.step()
        v = sim.volume()
        assert abs(v - prev) <= 1e-8 * v0, n
        assert sim.audit()["boundary_in"] == 0.0
        prev = v
    wc = interior(sim.get(abi.F_WC))
    assert wc[0, 0, 19] > 0.2 and wc[0, 0, 0] < 0.2


@pytest.mark.parametrize("backend", BACKENDS)
def test_audit_identity_with_compressible_storage(backend):
    # test_gw.cpp:168-186
    sim = make(backend, gw_base_case(1, 1, 20, 0.05, 0.5, specific_storage=1e-4))
    prev = sim.volume()
    for n in range(30):
        sim.step()
        v, res = closure(sim, prev)
        assert abs(res) <= 1e-12 * max(v, 1.0), n
        prev = v


@pytest.mark.parametrize("backend", BACKENDS)
def test_prescribed_head_top_infiltrates(backend):
    # test_gw.cpp:192-212
    c = gw_base_case(1, 1, 10, 0.05, 0.5, ic_const=0.06)
    c.set_bc("top", abi.BC_HEAD, 0.0)
    sim = make(backend, c)
    v0 = sim.volume()
    inflow = []
    run_steps(sim, 40, lambda s, dt: inflow.append(s.audit()["boundary_in"]))
    assert sum(inflow) > 0.0
    assert sim.volume() > v0
    assert interior(sim.get(abi.F_WC))[0, 0, 0] > 0.12


@pytest.mark.parametrize("backend", BACKENDS)
def test_prescribed_flux_top_delivers_exact_volume(backend):
    # test_gw.cpp:214-229 (negative = downward into the ground)
    c = gw_base_case(1, 1, 10, 0.1, 0.5)
    c.set_bc("top", abi.BC_FLUX, -1.0e-6)
    sim = make(backend, c)
    v0 = sim.volume()
    elapsed = run_steps(sim, 40)
    expected = 1.0e-6 * elapsed
    assert abs(sim.volume() - v0 - expected) <= 1e-12 * max(expected, 1.0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_gravity_bottom_drains_at_conductivity_rate(backend):
    # test_gw.cpp:231-264
    c = gw_base_case(1, 1, 10, 0.1, 0.5, specific_storage=1e-5, ic_form="water_table",
                     ic_const=0.0)
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)
    sim = make(backend, c)
    v0 = sim.volume()
    dtg = sim.step()
    a = sim.audit()
    assert abs(a["boundary_in"] - (-1.0e-5 * dtg)) <= 1e-18
    v, res = closure(sim, v0)
    assert abs(res) <= 1e-12
    assert v < v0


@pytest.mark.parametrize("backend", BACKENDS)
def test_prescribed_flux_bottom_injects(backend):
    # test_gw.cpp:266-285 (positive = upward = into the domain)
    c = gw_base_case(1, 1, 10, 0.1, 0.5)
    c.set_bc("bot", abi.BC_FLUX, 2.0e-6)
    sim = make(backend, c)
    v0 = sim.volume()
    elapsed = run_steps(sim, 20)
    assert abs(sim.volume() - v0 - 2.0e-6 * elapsed) <= 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_hydrostatic_side_holds_equilibrium(backend):
    # test_gw.cpp:287-321: middle cell of the west edge at the IC stage
    c = gw_base_case(3, 3, 10, 0.1, 0.5, ic_form="water_table", ic_const=-0.35)
    mask = np.zeros((3, 3), bool)
    mask[1, 0] = True
    c.set_bc("xm", abi.BC_HEAD_HYDROSTATIC, -0.35, mask)
    sim = make(backend, c)
    h0 = interior(sim.get(abi.F_HEAD)).copy()
    run_steps(sim, 10)
    assert np.max(np.abs(interior(sim.get(abi.F_HEAD)) - h0)) <= 1e-6


@pytest.mark.parametrize("backend", BACKENDS)
def test_side_flux_injects_with_inflow_positive(backend):
    # test_gw.cpp:323-345
    c = gw_base_case(3, 3, 5, 0.1, 0.5)
    mask = np.zeros((3, 3), bool)
    mask[1, 0] = True
    c.set_bc("xm", abi.BC_FLUX, 4.0e-6, mask)
    sim = make(backend, c)
    v0 = sim.volume()
    elapsed = run_steps(sim, 20)
    assert abs(sim.volume() - v0 - 4.0e-6 * 0.5 * elapsed) <= 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("full3d", [False, True])
def test_use_full3d_gates_lateral_flow(backend, full3d):
    # test_gw.cpp:351-373: x flux per area at (j=1, i=1, k=1) after one step
    moist = np.zeros((1, 2, 4))
    moist[0, 0, :] = 0.10
    moist[0, 1, :] = 0.30
    c = gw_base_case(2, 1, 4, 0.1, 0.5, use_full3d=full3d, ic_values=moist)
    sim = make(backend, c)
    sim.step()
    q = sim.get(abi.F_QX)[1, 1, 1] / sim.get(abi.F_AREA_X)[1, 1, 1]
    assert (q != 0.0) if full3d else (q == 0.0)


def _sand_over_clay(**kw):
    ids = np.array([0, 0, 0, 1, 1, 1, 1, 1, 1, 1])
    sand = dict(ks=3.0e-4, theta_s=0.46, theta_r=0.04, vg_alpha=5.9, vg_n=2.68, aev=-0.02)
    clay = dict(ks=1.52e-6, theta_s=0.4686, theta_r=0.106, vg_alpha=1.04, vg_n=1.395, aev=0.0)
    pick = lambda key: np.where(ids == 0, sand[key], clay[key]).reshape(1, 1, 10)
    c = gw_base_case(1, 1, 10, 0.05, 0.5, ic_const=0.14)
    c.ksx = c.ksy = c.ksz = pick("ks")
    for key in ("theta_s", "theta_r", "vg_alpha", "vg_n", "aev"):
        setattr(c, key, pick(key))
    for k, v in kw.items():
        setattr(c, k, v)
    c.set_bc("top", abi.BC_HEAD, 0.0)
    return c


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("pc", ["legacy", "consistent"])
@pytest.mark.parametrize("surplus", ["drop", "redistribute"])
def test_reallocation_surplus_modes_close_the_audit(backend, pc, surplus):
    # test_gw.cpp:375-434.  The reference runs it under predictor_corrector:
    # legacy, whose one-sided clay Ks on the sand/clay faces leaves corrector
    # moisture above theta(h) behind the front: drop discards it (more than
    # 1e-5 m^3 over 200 steps), redistribute keeps it in the column.  Under
    # the consistent pair only the closure identity is asserted.
    sim = make(backend, _sand_over_clay(reallocation_surplus=surplus, predictor_corrector=pc))
    v0 = prev = sim.volume()
    dropped = 0.0
    for n in range(200):
        sim.step()
        v, res = closure(sim, prev)
        assert abs(res) <= 1e-12, n
        dropped += sim.audit()["realloc_dropped"]
        prev = v
    assert sim.volume() > v0
    if pc == "legacy":
        if surplus == "drop":
            assert dropped > 1e-5
        else:
            assert dropped <= 1e-9


@pytest.mark.parametrize("backend", BACKENDS)
def test_reallocation_off_conserves_exactly(backend):
    # test_gw.cpp:436-478
    sim = make(backend, _sand_over_clay(reallocation="disabled"))
    v0 = prev = sim.volume()
    for n in range(200):
        sim.step()
        a = sim.audit()
        assert a["realloc_adjust"] == 0.0 and a["realloc_dropped"] == 0.0 and a["cpl_vent"] == 0.0
        v = sim.volume()
        assert abs((v - prev) - (a["boundary_in"] - a["ss_storage"] - a["vloss"])) <= 1e-12, n
        prev = v
    assert sim.volume() > v0


@pytest.mark.parametrize("backend", BACKENDS)
def test_reallocation_off_keeps_isolated_heads_on_curve(backend):
    # test_gw.cpp:480-504: h == h(theta) bitwise for theta < 0.9999 theta_s
    from oracle import oracle
    c = gw_base_case(1, 1, 10, 0.05, 0.5, reallocation="disabled")
    c.set_bc("top", abi.BC_FLUX, -1.0e-6)
    sim = make(backend, c)
    soil = (5.9, 2.68, 0.46, 0.04, -0.02)
    for n in range(20):
        sim.step()
        wc = interior(sim.get(abi.F_WC))[0, 0]
        h = interior(sim.get(abi.F_HEAD))[0, 0]
        for k in range(10):
            if wc[k] < 0.9999 * 0.46:
                ref = oracle.vg("head", soil, wc[k])
                if backend == "oracle":
                    assert h[k] == ref, (n, k)
                else:  # device pow differs from libm by at most a few ulp
                    assert abs(h[k] - ref) <= 1e-13 * abs(ref), (n, k)


@pytest.mark.parametrize("backend", BACKENDS)
def test_adaptive_step_grows_and_clamps(backend):
    # test_gw.cpp:506-528
    c = gw_base_case(1, 1, 8, 0.1, 0.1, specific_storage=1e-5, ic_form="water_table",
                     ic_const=0.0)
    c.dt_init, c.dt_min, c.dt_max = 0.1, 0.1, 1.0
    sim = make(backend, c)
    assert sim.next_dt() == 0.1
    sim.step()
    assert abs(sim.next_dt() - 0.125) <= 1e-12
    dtg = 0.0
    for _ in range(30):
        dtg = sim.step()
        assert 0.1 <= sim.next_dt() <= 1.0
    assert dtg == 1.0


@pytest.mark.parametrize("backend", BACKENDS)
def test_reallocate_restores_isolated_head_consistency(backend):
    # test_gw.cpp:530-545
    from oracle import oracle
    sim = make(backend, gw_base_case(1, 1, 10, 0.05, 0.5))
    sim.step()
    soil = (5.9, 2.68, 0.46, 0.04, -0.02)
    wc = interior(sim.get(abi.F_WC))[0, 0]
    h = interior(sim.get(abi.F_HEAD))[0, 0]
    for k in range(2, 8):
        assert abs(h[k] - oracle.vg("head", soil, wc[k])) <= 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("pos", [("west", 1, 0), ("east", 1, 2), ("south", 0, 1), ("north", 2, 1)])
def test_fully_masked_edge_column_conserves(backend, pos):
    # test_gw.cpp:547-619
    name, gj, gi = pos
    bath = np.zeros((3, 3))
    bath[gj, gi] = -1.0
    c = gw_base_case(3, 3, 4, 0.05, 0.5, bath=bath)
    sim = make(backend, c)
    assert sim.active_cells() == 32
    wc0 = sim.get(abi.F_WC)[gj + 1, gi + 1].copy()
    h0 = sim.get(abi.F_HEAD)[gj + 1, gi + 1].copy()
    v0 = prev = sim.volume()
    for n in range(30):
        sim.step()
        v, res = closure(sim, prev)
        assert abs(res) <= 1e-12, (name, n)
        assert sim.audit()["boundary_in"] == 0.0
        assert abs(v - prev) <= 1e-8 * v0
        prev = v
    assert np.array_equal(sim.get(abi.F_WC)[gj + 1, gi + 1], wc0)
    assert np.array_equal(sim.get(abi.F_HEAD)[gj + 1, gi + 1], h0)
    wc = sim.get(abi.F_WC)
    assert wc[2, 2, 3] > 0.2 and wc[2, 2, 0] < 0.2


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("harmonic", [False, True])
def test_saturated_heterogeneous_column_restore_books_nothing(backend, harmonic):
    # test_gw.cpp:833-885 (K1): consistent pair, head BCs top and bottom
    ids = [0, 3, 1, 4, 2, 5]
    ks = np.array([10.0 ** (-6.5 + 0.6 * n) for n in ids]).reshape(1, 1, 6)
    c = gw_base_case(1, 1, 6, 0.25, 10.0, ks=ks, ic_form="water_table", ic_const=5.0,
                     face_conductivity="harmonic" if harmonic else "arithmetic")
    c.set_bc("top", abi.BC_HEAD, 2.0)
    c.set_bc("bot", abi.BC_HEAD, 3.3)
    sim = make(backend, c)
    through = churn = 0.0
    for _ in range(20):
        dtg = sim.step()
        a = sim.audit()
        through += abs(sim.get(abi.F_QZF)[1, 1, 3]) * dtg
        churn += abs(a["realloc_adjust"]) + 2.0 * a["realloc_dropped"]
    assert through > 1e-5
    assert churn <= 1e-10 * through


@pytest.mark.parametrize("backend", BACKENDS)
def test_saturated_heterogeneous_column_legacy_churns(backend):
    # test_gw.cpp:885-927: the same stack under predictor_corrector: legacy
    # books the corrector/predictor gap (the red control of the test above)
    ids = [0, 3, 1, 4, 2, 5]
    ks = np.array([10.0 ** (-6.5 + 0.6 * n) for n in ids]).reshape(1, 1, 6)
    c = gw_base_case(1, 1, 6, 0.25, 10.0, ks=ks, ic_form="water_table", ic_const=5.0,
                     predictor_corrector="legacy")
    c.set_bc("top", abi.BC_HEAD, 2.0)
    c.set_bc("bot", abi.BC_HEAD, 3.3)
    sim = make(backend, c)
    through = churn = 0.0
    for _ in range(20):
        dtg = sim.step()
        a = sim.audit()
        through += abs(sim.get(abi.F_QZF)[1, 1, 3]) * dtg
        churn += abs(a["realloc_adjust"]) + 2.0 * a["realloc_dropped"]
    assert through > 1e-5
    assert churn > 1e-4 * through


@pytest.mark.parametrize("backend", BACKENDS)
def test_homogeneous_consistent_is_bitwise_legacy(backend):
    # test_gw.cpp:929-955 (K3): homogeneous soil, uncoupled: every face rule
    # of the two pairs coincides, so the heads agree bitwise
    heads = []
    for pc in ("legacy", "consistent"):
        c = gw_base_case(2, 1, 4, 0.25, 10.0, predictor_corrector=pc)
        c.set_bc("top", abi.BC_HEAD, 0.1)
        sim = make(backend, c)
        for _ in range(30):
            sim.step()
        heads.append(sim.get(abi.F_HEAD)[1, 1, :].copy())
    assert np.array_equal(heads[0], heads[1])


@pytest.mark.parametrize("backend", BACKENDS)
def test_ponded_top_face_consistent_is_bitwise_legacy(backend):
    # test_gw.cpp:957-987 (K2): the uncoupled head top was Ksz in both pairs
    thetas = []
    for pc in ("legacy", "consistent"):
        c = gw_base_case(1, 1, 6, 0.2, 5.0, ic_const=0.15, predictor_corrector=pc)
        c.set_bc("top", abi.BC_HEAD, 0.3)
        sim = make(backend, c)
        for _ in range(50):
            sim.step()
        thetas.append(sim.get(abi.F_WC)[1, 1, :].copy())
    assert np.array_equal(thetas[0], thetas[1])


def _saturated_pair(pc):
    # two saturated cells in x (Ks 1e-5 and 4e-5) between head sides, no
    # storage, nothing after the fluxes that moves h (reallocation disabled)
    ks = np.array([1e-5, 4e-5]).reshape(1, 2, 1)
    c = gw_base_case(2, 1, 1, 0.5, 10.0, ks=ks, ic_form="water_table", ic_const=5.0,
                     reallocation="disabled", predictor_corrector=pc)
    c.set_bc("xm", abi.BC_HEAD, 2.0)
    c.set_bc("xp", abi.BC_HEAD, 1.0)
    return c


@pytest.mark.parametrize("backend", BACKENDS)
def test_legacy_lateral_face_takes_the_minus_cells_ks(backend):
    # Corrector.cpp:212-236: with both cells saturated (K = Ks) the legacy
    # interior face conducts at the minus cell's Ks, the consistent one at
    # the flanks' mean; the face flux over the head difference then differs
    # by exactly that ratio (area and spacing cancel)
    rate = {}
    for pc in ("legacy", "consistent"):
        sim = make(backend, _saturated_pair(pc))
        sim.step()
        h = sim.get(abi.F_HEAD)[1, 1:3, 0]
        assert np.all(h > 0) and h[0] != h[1]
        rate[pc] = sim.get(abi.F_QX)[1, 1, 0] / (h[1] - h[0])
    assert rate["legacy"] / rate["consistent"] == pytest.approx(1e-5 / (0.5 * (1e-5 + 4e-5)),
                                                                rel=1e-12)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("pc", ["legacy", "consistent"])
def test_coupled_wet_top_face_conductivity(backend, pc):
    # Predictor.cpp:293-302: a wet coupled top conducts at Ksz (consistent)
    # or 0.5 (Ksz + K(h_top)) (legacy); the corrector's face array holds it
    from oracle.oracle import vg
    ks = 2e-5
    c = gw_base_case(1, 1, 8, 0.1, 30.0, ks=ks, ic_const=0.15, reallocation="disabled",
                     predictor_corrector=pc)
    dept = np.zeros((3, 3))
    dept[1, 1] = 0.05
    c.coupling = dict(eta=dept.copy(), dept=dept.copy(), seep_accum=np.zeros_like(dept),
                      avail=dept.copy(), gain=np.zeros_like(dept), bounce=np.zeros_like(dept),
                      min_depth=1e-4, vent_all=0)
    sim = make(backend, c)
    sim.step()
    h_top = sim.get(abi.F_HEAD)[1, 1, 0]
    assert h_top < 0  # unsaturated, so the two rules differ
    soil = (TEST_SOIL["vg_alpha"], TEST_SOIL["vg_n"], TEST_SOIL["theta_s"],
            TEST_SOIL["theta_r"], TEST_SOIL["aev"])
    want = 0.5 * (ks + vg("conductivity", soil, h_top, ks)) if pc == "legacy" else ks
    assert sim.get(abi.F_KZF)[1, 1, 0] == pytest.approx(want, rel=1e-14)


@pytest.mark.parametrize("backend", BACKENDS)
def test_sink_in_saturated_confined_cell_drops_head_by_sink_over_ss(backend):
    # test_gw.cpp:780-823: one sealed saturated cell, Ss 1e-4, a constant
    # sink S = 1e-6 / 0.5 for one step: h - hn = -S dtg / Ss (1e-12) and
    # theta stays theta_s (1e-14).  The sink is given directly (the
    # vegetation module that derives it from T_pot is not part of the step).
    c = gw_base_case(1, 1, 1, 0.5, 60.0, ks=0.0, specific_storage=1e-4, reallocation="disabled",
                     ic_form="head", ic_const=2.0)
    sim = make(backend, c)
    sink = np.zeros(sim.field_shape(abi.F_SINK))
    rate = 1e-6 / 0.5
    sink[1, 1, 0] = rate
    sim.set_sink(sink)
    hn = sim.get(abi.F_HEAD)[1, 1, 0]
    sim.step()
    h = sim.get(abi.F_HEAD)[1, 1, 0]
    expected = -rate * 60.0 / 1e-4
    assert (h - hn) / expected == pytest.approx(1.0, abs=1e-12)
    assert sim.get(abi.F_WC)[1, 1, 0] == pytest.approx(TEST_SOIL["theta_s"], abs=1e-14)
    assert sim.audit()["tr_applied"] == pytest.approx(rate * 0.5 * 60.0, rel=1e-15)


def _sink_column():
    c = gw_base_case(2, 2, 10, 0.1, 30.0, specific_storage=1e-5)
    c.set_bc("top", abi.BC_FLUX, -1e-6)
    sink = np.zeros((4, 4, 10))
    sink[1:3, 1:3, :5] = np.linspace(1e-7, 5e-7, 5)
    return c, sink


@pytest.mark.parametrize("backend", BACKENDS)
def test_sink_audit_closes(backend):
    # the coupled closure's groundwater part with the transpiration channel:
    # dV = boundary_in - ss_storage + realloc_adjust - vloss - tr_applied,
    # and tr_applied = sum S V dtg (the sink leaves through no face)
    c, sink = _sink_column()
    sim = make(backend, c)
    sim.set_sink(sink)
    prev = sim.volume()
    for n in range(30):
        sim.step()
        v, res = closure(sim, prev)
        a = sim.audit()
        assert a["tr_applied"] == pytest.approx(sink.sum() * 0.1 * 30.0, rel=1e-13)
        assert abs(res + a["tr_applied"]) <= 1e-12, n
        prev = v


@pytest.mark.parametrize("backend", BACKENDS)
def test_groundwater_loop_residual_closes_with_a_sink(backend):
    # run_groundwater_loop books the transpiration channel in its per-step
    # mass residual (Simulation.cpp:1082-1126 with trApplied)
    from frehg2_amd.config import run_groundwater_loop
    c, sink = _sink_column()
    sim = make(backend, c)
    sim.set_sink(sink)
    recs = run_groundwater_loop(sim, None, t_end=30.0 * 12)
    assert len(recs) == 12
    for r in recs:
        assert r["tr_applied"] > 0.0
        assert abs(r["residual"]) <= 1e-12, r


@pytest.mark.parametrize("backend", BACKENDS)
def test_detached_sink_is_no_sink(backend):
    c, sink = _sink_column()
    ref = make(backend, c)
    sim = make(backend, c)
    sim.set_sink(sink)
    sim.set_sink(None)
    for _ in range(5):
        ref.step()
        sim.step()
    assert np.array_equal(ref.get(abi.F_HEAD), sim.get(abi.F_HEAD))
    assert np.array_equal(ref.get(abi.F_WC), sim.get(abi.F_WC))
    assert sim.audit()["tr_applied"] == 0.0


def _density_column(ratios, gradient_ratios):
    """A closed saturated 1x1x10 column whose head rises downward by the face
    mean of `gradient_ratios` per spacing (hydrostatic for those ratios)."""
    nz, dz = 10, 0.1
    h = np.empty(nz)
    h[0] = 0.3
    for k in range(1, nz):
        h[k] = h[k - 1] + dz * 0.5 * (gradient_ratios[k - 1] + gradient_ratios[k])
    c = gw_base_case(1, 1, nz, dz, 60.0, specific_storage=1e-5, ic_form="head",
                     ic_values=h.reshape(1, 1, nz))
    r = np.broadcast_to(ratios, (3, 3, nz)).copy()
    return c, r


@pytest.mark.parametrize("backend", BACKENDS)
def test_density_hydrostatic_column_stays_at_rest(backend):
    # Baroclinic.cpp face means + Corrector.cpp:502-525: a z face carries no
    # flux when dh/dz equals its density ratio, the mean of its two cells; a
    # column built on that rule stays at rest under any viscosity ratio,
    # and one built on the freshwater gradient does not
    rng = np.random.default_rng(7)
    ratios = 1.0 + 0.05 * rng.random(10)
    visc = 0.9 + 0.1 * rng.random((3, 3, 10))
    c, r = _density_column(ratios, ratios)
    sim = make(backend, c)
    sim.set_density(r, visc)
    h0 = sim.get(abi.F_HEAD)[1, 1, :].copy()
    for _ in range(10):
        sim.step()
    assert np.max(np.abs(sim.get(abi.F_HEAD)[1, 1, :] - h0)) <= 1e-10
    assert np.max(np.abs(sim.get(abi.F_QZF)[1, 1, 1:10])) <= 1e-18
    c, r = _density_column(ratios, np.ones(10))
    sim = make(backend, c)
    sim.set_density(r, visc)
    h0 = sim.get(abi.F_HEAD)[1, 1, :].copy()
    sim.step()
    assert np.max(np.abs(sim.get(abi.F_HEAD)[1, 1, :] - h0)) > 1e-6


@pytest.mark.parametrize("backend", BACKENDS)
def test_uniform_viscosity_ratio_scales_conductivity(backend):
    # K enters every Darcy term as K v: a uniform r_visc = c is Ks scaled by c
    # wherever no face keeps the default 1 (flux top, no-flux bottom and sides)
    cv = 0.8

    def case(ks):
        c = gw_base_case(2, 2, 10, 0.1, 30.0, ks=ks, specific_storage=1e-5)
        c.set_bc("top", abi.BC_FLUX, -1e-6)
        return c
    a = make(backend, case(1e-5))
    a.set_density(np.ones((4, 4, 10)), np.full((4, 4, 10), cv))
    b = make(backend, case(1e-5 * cv))
    for _ in range(20):
        a.step()
        b.step()
    ha, hb = a.get(abi.F_HEAD)[1:-1, 1:-1], b.get(abi.F_HEAD)[1:-1, 1:-1]
    assert np.max(np.abs(ha - hb)) <= 1e-10 * np.max(np.abs(hb))
    plain = make(backend, case(1e-5))
    for _ in range(20):
        plain.step()
    assert np.max(np.abs(plain.get(abi.F_HEAD)[1:-1, 1:-1] - ha)) > 1e-6  # the ratio mattered


@pytest.mark.parametrize("backend", BACKENDS)
def test_unit_ratios_are_no_ratios(backend):
    c, sink = _sink_column()
    ref = make(backend, c)
    sim = make(backend, c)
    sim.set_density(np.ones((4, 4, 10)), np.ones((4, 4, 10)), np.ones((4, 4)), np.ones((4, 4)))
    for _ in range(5):
        ref.step()
        sim.step()
    assert np.array_equal(ref.get(abi.F_HEAD), sim.get(abi.F_HEAD))
    assert np.array_equal(ref.get(abi.F_WC), sim.get(abi.F_WC))


@pytest.mark.parametrize("backend", BACKENDS)
def test_et_channel_replaces_the_top_condition(backend):
    # Corrector.cpp:471-472 and Predictor.cpp:675-679: an uncoupled channel
    # column's top face carries exactly qVap A, guard-free, whatever its
    # configured top flux; the others keep theirs.  The audit closes with it
    # in boundary_in.
    c = gw_base_case(2, 1, 10, 0.1, 30.0, specific_storage=1e-5)
    c.set_bc("top", abi.BC_FLUX, -1e-6)
    sim = make(backend, c)
    cols = np.zeros((3, 4), dtype=np.int32)
    cols[1, 1] = 1
    qvap = np.zeros((3, 4))
    qvap[1, 1] = 3e-7
    sim.set_et(cols, qvap)
    prev = sim.volume()
    for n in range(10):
        sim.step()
        q = sim.get(abi.F_QZF)
        assert q[1, 1, 0] == 3e-7 * 1.0 and q[1, 2, 0] == -1e-6 * 1.0
        v, res = closure(sim, prev)
        assert abs(res) <= 1e-12
        assert sim.audit()["boundary_in"] == pytest.approx((1e-6 - 3e-7) * 30.0, rel=1e-12)
        prev = v


@pytest.mark.parametrize("backend", BACKENDS)
def test_et_channel_books_evaporation_in_coupled_runs(backend):
    # Corrector.cpp:609-615: the channel's volume leaves the seepage
    # accumulator and is booked as evaporation, any sign, no depth test
    c = gw_base_case(1, 1, 8, 0.1, 30.0, ic_const=0.3)
    z = np.zeros((3, 3))
    c.coupling = dict(eta=z.copy(), dept=z.copy(), seep_accum=z.copy(), avail=z.copy(),
                      gain=z.copy(), bounce=z.copy(), min_depth=1e-4, vent_all=0)
    sim = make(backend, c)
    sim.set_et(np.ones((3, 3), dtype=np.int32), np.full((3, 3), 2e-7))
    sim.step()
    qtop = sim.get(abi.F_QZF)[1, 1, 0]
    assert sim.audit()["cpl_evap"] == pytest.approx(2e-7 * 30.0, rel=1e-14)
    assert sim.coupling_state("seep_accum")[1, 1] == pytest.approx(qtop * 30.0 - 2e-7 * 30.0,
                                                                   rel=1e-12)
