"""YAML keys -> GwCase (frehg2_amd.config) and the Warrick infiltration pin.

The deck below uses the reference's keys with the values of its
re-1d-infiltration validation case (1D column, 100 x 0.01 m, Warrick soil,
ponded top).  Warrick's semi-analytical wetting-front depths are the data the
reference keeps beside that case (validation/re-1d-infiltration/makeplot.py:
66-71, percent of the column at 11700, 23400 and 46800 s for theta = 0.0825,
0.165, 0.2475); its regression gate holds the theta = 0.165 front within 5 %
(tests/regression/tolerances/b2-gw.yaml:19-24).
"""
import numpy as np
import pytest

from frehg2_amd import abi
from frehg2_amd.config import (TimeSeries, coupled_density_ratios, density_ratios, load_config, polygon_contains,
                               run_groundwater_loop)

DECK = {
    "simulation": {"id": "re-1d-infiltration"},
    "domain": {"nx": 1, "ny": 1, "nz": 100, "dx": 1.0, "dy": 1.0, "dz": 0.01,
               "dz_stretch": 1.0, "bottom_elevation": {"constant": 0.0},
               "follow_terrain": False},
    "groundwater": {"scheme": "pca", "use_full3d": True,
                    "timestep": {"dt_init": 1e-4, "dt_min": 1e-4, "dt_max": 0.4,
                                 "courant_max": 2.0},
                    "specific_storage": 0.0, "reallocation_surplus": "redistribute"},
    "soil": {"types": [{"name": "warrick", "ksx": 2.89e-6, "ksy": 2.89e-6, "ksz": 2.89e-6,
                        "theta_s": 0.33, "theta_r": 0.0, "vg_alpha": 1.43, "vg_n": 1.56,
                        "aev": 0.0}],
             "map": {"constant": "warrick"}},
    "initial_conditions": {"groundwater": {"moisture": {"constant": 0.03}}},
    "boundary_conditions": [
        {"name": "ponded-top", "region": {"polygon": [[-0.1, -0.1], [1.1, -0.1], [1.1, 1.1],
                                                      [-0.1, 1.1]]},
         "target": "groundwater_top", "kind": "head", "value": {"constant": 1.0e-5}}],
}
WARRICK_T = [11700.0, 23400.0, 46800.0]
WARRICK_PCT = {0.0825: [74.58, 61.32, 39.02], 0.165: [75.20, 62.13, 40.04],
               0.2475: [77.12, 64.54, 42.80]}


def front_depth(wc_col, dz, level):
    """Depth below the surface where theta first drops below `level`."""
    zc = (np.arange(wc_col.size) + 0.5) * dz
    k = int(np.argmax(wc_col < level))
    if k == 0:
        return 0.0
    w0, w1 = wc_col[k - 1], wc_col[k]
    return zc[k - 1] + (w0 - level) / (w0 - w1) * (zc[k] - zc[k - 1])


def test_deck_maps_to_case():
    case, sched = load_config(DECK)
    assert (case.nx, case.ny, case.nz) == (1, 1, 100)
    assert case.ic_form == "moisture" and case.ic_const == 0.03
    assert case.bc["top"][0][0, 0] == abi.BC_HEAD and case.bc["top"][1][0, 0] == 1e-5
    assert np.all(case.soil_array("vg_n") == 1.56)
    assert case.reallocation_surplus == "redistribute"
    assert not sched.time_dependent
    assert case.predictor_corrector == "consistent"


def test_scheme_keys():
    import copy
    deck = copy.deepcopy(DECK)
    deck["groundwater"]["predictor_corrector"] = "legacy"
    case, _ = load_config(deck)
    assert case.predictor_corrector == "legacy" and case.to_desc()[0].predictor_corrector == 0
    # saturated_contact: legacy reads cells its own sweep is rewriting
    # (Reallocate.cpp:177): its result depends on the loop order
    deck["groundwater"]["density_coupling"] = {"enabled": True, "reallocation_gradient": "density"}
    case, _ = load_config(deck)
    dc = case.density_coupling
    assert dc["beta_saline"] == 0.000744 and dc["reallocation_gradient"] == "density"
    # sea water at 35 PSU (Baroclinic.cpp:78-86 with the default betas)
    r_rho, r_visc = density_ratios(dc, salinity=np.array([35.0]), temperature=np.array([20.0]))
    assert r_rho[0] == pytest.approx(1.02604, rel=1e-15)
    assert r_visc[0] == pytest.approx(1.0 / 1.077, rel=1e-15)
    deck["groundwater"]["saturated_contact"] = "legacy"
    with pytest.raises(ValueError):
        load_config(deck)


def test_coupled_density_ratios_take_the_top_face_from_the_surface():
    """Coupled runs take the top face's ghost ratio from the surface scalars
    (Baroclinic.cpp:172-180): the helper hands them to set_density, and they
    change the top face's flux against the uncoupled own-cell rule."""
    from frehg2_amd import cases
    from oracle.oracle import OracleRichards
    dc = {"beta_saline": 0.000744, "thermal_expansion": 0.00021, "reference_temperature": 20.0,
          "beta_saline_viscosity": 0.0022, "reallocation_gradient": "freshwater"}
    c = cases.coupled_128x128x64(member=0, nx=6, ny=5, nz=8)
    shape3 = (c.ny + 2, c.nx + 2, c.nz)
    shape2 = (c.ny + 2, c.nx + 2)
    s3 = np.full(shape3, 5.0)
    s2 = np.full(shape2, 30.0)
    kw = coupled_density_ratios(dc, salinity=s3, surface_salinity=s2)
    assert kw["r_rho_top"].shape == shape2
    assert kw["r_rho_top"][0, 0] == pytest.approx(1.0 + 0.000744 * 30.0, rel=1e-15)
    assert kw["r_visc_top"][0, 0] == pytest.approx(1.0 / (1.0 + 0.0022 * 30.0), rel=1e-15)
    with pytest.raises(ValueError):
        coupled_density_ratios(dc, salinity=s3)
    runs = []
    for top in (True, False):
        o = OracleRichards(c)
        if top:
            o.set_density(**kw)
        else:
            o.set_density(kw["r_rho"], kw["r_visc"], reallocation_gradient="freshwater")
        o.step(c.dt)
        runs.append(o.get(abi.F_QZF)[1:-1, 1:-1, 0])  # the top faces
    assert np.max(np.abs(runs[0] - runs[1])) > 0.0


def test_reference_preconditioners_map_to_zline_visibly():
    import copy
    import warnings
    deck = copy.deepcopy(DECK)
    deck.setdefault("solver", {}).setdefault("groundwater", {})["preconditioner"] = "amg"
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        case, _ = load_config(deck)
    assert case.preconditioner == "zline" and case.preconditioner_requested == "amg"
    assert any("amg" in str(x.message) for x in w)
    deck["solver"]["groundwater"]["preconditioner"] = "jacobi"
    case, _ = load_config(deck)
    assert case.preconditioner == "jacobi"
    deck["solver"]["groundwater"]["preconditioner"] = "ilu-magic"
    with pytest.raises(ValueError):
        load_config(deck)


def test_polygon_and_series_semantics(tmp_path):
    sq = [(0, 0), (2, 0), (2, 2), (0, 2)]
    assert polygon_contains(sq, 1, 1) and polygon_contains(sq, 2, 1)  # edge counts as inside
    assert not polygon_contains(sq, 2.5, 1)
    f = tmp_path / "s.dat"
    f.write_text("# t v\n0 1.0\n10 3.0\n")
    s = TimeSeries.from_file(f)
    assert s.value(-5) == 1.0 and s.value(5) == 2.0 and s.value(50) == 3.0


def test_side_targets_select_edge_faces():
    deck = {k: v for k, v in DECK.items()}
    deck["domain"] = dict(DECK["domain"], nx=3, ny=2, nz=4)
    deck["boundary_conditions"] = [
        {"name": "w", "region": {"polygon": [[-1, -1], [0.9, -1], [0.9, 0.9], [-1, 0.9]]},
         "target": "groundwater_side", "kind": "head", "value": {"hydrostatic": {"eta": -0.2}}}]
    case, _ = load_config(deck)
    # cell (j=0, i=0) is a west and a south edge cell
    assert case.bc["xm"][0][0, 0] == abi.BC_HEAD_HYDROSTATIC
    assert case.bc["ym"][0][0, 0] == abi.BC_HEAD_HYDROSTATIC
    assert case.bc["xm"][0][1, 0] == abi.BC_NOFLUX


@pytest.mark.parametrize("backend", ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def test_wetting_front_matches_warrick(backend):
    """The re-1d-infiltration deck end to end through the gw-only driver loop
    (Simulation.cpp:1672-1770): adaptive dt from 1e-4 s to 0.4 s, about
    1.2e5 steps to 46800 s, on the oracle and on the HIP path."""
    case, sched = load_config(DECK)
    case.rtol, case.atol = 1e-10, 1e-30
    if backend == "oracle":
        from oracle.oracle import OracleRichards
        o = OracleRichards(case)
    else:
        from frehg2_amd.gw import RichardsSolver
        o = RichardsSolver(case)
        print("[deck] re-1d-infiltration on the HIP path", flush=True)
    snaps = {}
    it = iter(WARRICK_T)
    target = [next(it)]

    count = [0]

    def grab(sim, rec):
        count[0] += 1
        if count[0] % 20000 == 0:
            print(f"[deck] {backend} step {count[0]} t={rec['t']:.0f} s", flush=True)
        if target[0] is not None and rec["t"] >= target[0]:
            snaps[target[0]] = sim.get(abi.F_WC)[1, 1].copy()
            target[0] = next(it, None)

    recs = run_groundwater_loop(o, sched, WARRICK_T[-1], on_step=grab)
    assert max(abs(r["residual"]) for r in recs) <= 1e-12
    for n, t in enumerate(WARRICK_T):
        for level, pct in WARRICK_PCT.items():
            z_model = front_depth(snaps[t], 0.01, level)
            z_ref = 1.0 - pct[n] / 100.0
            rel = abs(z_model - z_ref) / z_ref
            print(f"[deck] {backend} t={t:.0f} theta={level}: model {z_model:.4f} m, "
                  f"Warrick {z_ref:.4f} m, rel {rel:.3f}", flush=True)
            if level == 0.165:
                assert rel <= 0.05, (t, level, z_model, z_ref)
