"""YAML configuration of the groundwater path, with the reference's keys.

Reads the subset of the reference schema this path consumes
(src/core/Config.cpp:60-80 domain, :305-341 groundwater, :370-394 soil,
:415-420 initial_conditions.groundwater, :487-520 boundary_conditions,
:810-827 solver.groundwater) and builds a ``GwCase``.  Two extensions,
absent from the reference schema, are accepted:

  soil.fields.{ksx,ksy,ksz,theta_s,theta_r,vg_alpha,vg_n,aev}: .npy files of
      per-cell values (ny, nx, nz) -- continuous soil fields;
  solver.groundwater.precision: fp64 | mixed.

Boundary polygons are rasterized on cell centres exactly as the reference's
BoundarySet does (src/bc/BoundarySet.cpp:45-146; on-edge points count as
inside, src/bc/Polygon.cpp): footprint targets select (j, i) columns, side
targets select the domain-edge faces of the selected edge cells.  Later
conditions overwrite earlier ones on shared cells, as the reference's marker
kernels do.  Series values follow TimeSeries::value (src/core/TimeSeries.cpp):
linear interpolation, clamped at both ends.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np
import yaml

from . import abi
from .cases import GwCase

SIDES = ("top", "bot", "xm", "xp", "ym", "yp")


class TimeSeries:
    """Piecewise-linear series of (time, value) samples, clamped at both ends."""

    def __init__(self, times, values):
        t = np.asarray(times, dtype=np.float64)
        v = np.asarray(values, dtype=np.float64)
        if t.size == 0 or t.size != v.size:
            raise ValueError("TimeSeries needs matching, non-empty time and value lists")
        if np.any(np.diff(t) <= 0):
            raise ValueError("TimeSeries times must be strictly increasing")
        self.t, self.v = t, v

    @classmethod
    def from_file(cls, path):
        ts, vs = [], []
        for n, line in enumerate(Path(path).read_text().splitlines(), 1):
            line = line.split("#", 1)[0].split()
            if not line:
                continue
            if len(line) != 2:
                raise ValueError(f"{path} line {n}: expected 'time value'")
            ts.append(float(line[0]))
            vs.append(float(line[1]))
        return cls(ts, vs)

    def value(self, t: float) -> float:
        if t <= self.t[0]:
            return float(self.v[0])
        if t >= self.t[-1]:
            return float(self.v[-1])
        hi = int(np.searchsorted(self.t, t, side="right"))
        lo = hi - 1
        w = (t - self.t[lo]) / (self.t[hi] - self.t[lo])
        return float(self.v[lo] + w * (self.v[hi] - self.v[lo]))


def _on_segment(a, b, x, y):
    cross = (b[0] - a[0]) * (y - a[1]) - (b[1] - a[1]) * (x - a[0])
    scale = max(abs(a[0]), abs(a[1]), abs(b[0]), abs(b[1]), abs(x), abs(y), 1.0)
    if abs(cross) > 1e-12 * scale * scale:
        return False
    dot = (x - a[0]) * (b[0] - a[0]) + (y - a[1]) * (b[1] - a[1])
    len2 = (b[0] - a[0]) ** 2 + (b[1] - a[1]) ** 2
    return -1e-12 * scale * scale <= dot <= len2 + 1e-12 * scale * scale


def polygon_contains(poly, x, y) -> bool:
    xs = [p[0] for p in poly]
    ys = [p[1] for p in poly]
    if x < min(xs) - 1e-12 or x > max(xs) + 1e-12 or y < min(ys) - 1e-12 or y > max(ys) + 1e-12:
        return False
    n = len(poly)
    for v in range(n):
        if _on_segment(poly[v], poly[(v + 1) % n], x, y):
            return True
    inside = False
    for v in range(n):
        a, b = poly[v], poly[(v + 1) % n]
        if (a[1] > y) != (b[1] > y):
            xc = a[0] + (y - a[1]) * (b[0] - a[0]) / (b[1] - a[1])
            if x < xc:
                inside = not inside
    return inside


@dataclass
class BcEntry:
    which: str
    code: int
    mask: np.ndarray
    constant: float = 0.0
    series: Optional[TimeSeries] = None

    def value(self, t):
        return self.series.value(t) if self.series is not None else self.constant


class BoundarySchedule:
    """Per-step boundary values (RichardsSolver::updateBoundaryValues)."""

    def __init__(self, entries, ny, nx):
        self.entries = entries
        self.ny, self.nx = ny, nx

    @property
    def time_dependent(self):
        return any(e.series is not None for e in self.entries)

    def values(self, t: float) -> dict:
        out = {}
        for e in self.entries:
            arr = out.setdefault(SIDES.index(e.which), np.zeros((self.ny, self.nx)))
            arr[e.mask] = e.value(t)
        return out


def _read_field(base: Path, spec, shape, name):
    if "constant" in spec:
        return float(spec["constant"])
    p = base / spec["file"]
    if p.suffix == ".npy":
        arr = np.load(p, allow_pickle=False)
    else:  # the reference's flat text list, (j*nx + i)*nz + k order
        arr = np.loadtxt(p, dtype=np.float64)
    arr = np.asarray(arr, dtype=np.float64)
    if arr.size != int(np.prod(shape)):
        raise ValueError(f"{name}: {arr.size} values for shape {shape}")
    return arr.reshape(shape)


# The reference's PETSc/hypre preconditioner names and what runs here.
PC_REMAP = {"bjacobi-icc": "zline", "amg": "zline", "gamg": "zline"}


def load_config(src, base_dir=None):
    """Build (GwCase, BoundarySchedule) from a YAML path or an already-parsed dict."""
    if isinstance(src, (str, Path)):
        base = Path(base_dir) if base_dir else Path(src).resolve().parent
        cfg = yaml.safe_load(Path(src).read_text())
    else:
        base = Path(base_dir or ".")
        cfg = src
    dom = cfg["domain"]
    nx, ny, nz = int(dom["nx"]), int(dom["ny"]), int(dom["nz"])
    dx, dy, dz = float(dom["dx"]), float(dom["dy"]), float(dom["dz"])
    bath = _read_field(base, dom.get("bottom_elevation", {"constant": 0.0}), (ny, nx), "bath")
    gw = cfg.get("groundwater", {})
    ts = gw.get("timestep", {})
    soil = cfg.get("soil", {})
    case = GwCase(name=cfg.get("simulation", {}).get("id", "yaml"), nx=nx, ny=ny, nz=nz,
                  dx=dx, dy=dy, dz=dz, dz_stretch=float(dom.get("dz_stretch", 1.0)),
                  follow_terrain=bool(dom.get("follow_terrain", False)),
                  terrain_layers=str(dom.get("terrain_layers", "scaled")))
    if isinstance(bath, float):
        case.bath_const = bath
    else:
        case.bath = bath
    case.predictor_corrector = str(gw.get("predictor_corrector", "consistent"))
    if case.predictor_corrector not in ("consistent", "legacy"):
        raise ValueError("groundwater.predictor_corrector: consistent or legacy")
    if str(gw.get("saturated_contact", "head")) != "head":
        raise ValueError("groundwater.saturated_contact: only 'head' is supported")
    case.use_full3d = bool(gw.get("use_full3d", True))
    case.specific_storage = float(gw.get("specific_storage", 0.0))
    case.face_conductivity = str(gw.get("face_conductivity", "arithmetic"))
    case.reallocation = str(gw.get("reallocation", "legacy"))
    case.reallocation_surplus = str(gw.get("reallocation_surplus", "drop"))
    dc = gw.get("density_coupling") or {}
    if bool(dc.get("enabled", False)):
        # src/core/Config.cpp:345-356; the transport module supplies the
        # scalars, density_ratios() turns them into the solver's input
        case.density_coupling = dict(
            beta_saline=float(dc.get("beta_saline", 0.000744)),
            beta_saline_viscosity=float(dc.get("beta_saline_viscosity", 0.0022)),
            thermal_expansion=float(dc.get("thermal_expansion", 0.0)),
            reference_temperature=float(dc.get("reference_temperature", 20.0)),
            reallocation_gradient=str(dc.get("reallocation_gradient", "freshwater")))
    case.dt_init = float(ts.get("dt_init", 0.0))
    case.dt_min = float(ts.get("dt_min", 0.0))
    case.dt_max = float(ts.get("dt_max", 0.0))
    case.dq_grow = float(ts.get("dq_grow", 0.01))
    case.dq_shrink = float(ts.get("dq_shrink", 0.02))
    case.courant_max = float(ts.get("courant_max", 2.0))
    case.dt = case.dt_init
    # soil: types x map, or per-cell fields (extension)
    keys = {"ksx": "ksx", "ksy": "ksy", "ksz": "ksz", "theta_s": "theta_s",
            "theta_r": "theta_r", "vg_alpha": "vg_alpha", "vg_n": "vg_n", "aev": "aev"}
    if "fields" in soil:
        for k in keys:
            v = _read_field(base, soil["fields"][k], (ny, nx, nz), k) if k in soil["fields"] \
                else (0.0 if k == "aev" else None)
            if v is None:
                raise ValueError(f"soil.fields.{k} is required")
            setattr(case, k, v)
    else:
        types = soil["types"]
        names = [t["name"] for t in types]
        m = soil["map"]
        if "constant" in m:
            idx = np.full((ny, nx, nz), names.index(m["constant"]), dtype=np.int64)
        else:
            raw = _read_field(base, {"file": m["file"]}, (ny, nx, nz), "soil.map")
            idx = np.rint(raw).astype(np.int64)
            if np.any(np.abs(raw - idx) > 1e-12) or idx.min() < 0 or idx.max() >= len(types):
                raise ValueError("soil.map.file holds an invalid type index")
        for k in keys:
            table = np.array([float(t.get(k, 0.0)) for t in types])
            setattr(case, k, table[idx])
    # initial conditions
    ic = cfg.get("initial_conditions", {}).get("groundwater", {})
    for form in ("moisture", "head", "water_table"):
        if form in ic:
            case.ic_form = form
            shape = (ny, nx) if form == "water_table" else (ny, nx, nz)
            v = _read_field(base, ic[form], shape, form)
            if isinstance(v, float):
                case.ic_const, case.ic_values = v, None
            else:
                case.ic_values = v
            break
    # solver
    sg = cfg.get("solver", {}).get("groundwater", {})
    case.rtol = float(sg.get("rtol", 1e-8))
    case.atol = float(sg.get("atol", 1e-14))
    case.max_iterations = int(sg.get("max_iterations", 1000))
    pc = str(sg.get("preconditioner", "bjacobi-icc"))
    # The reference's preconditioners (LinearSystem.cpp:158-221) run on
    # PETSc/hypre; this path's PCG offers point Jacobi, the z-line block
    # Jacobi and (mixed form, extension key) a Jacobi-Chebyshev polynomial.
    # bjacobi-icc (the reference default), amg and gamg map to z-line: the
    # remap is recorded on the case and announced.
    if pc in PC_REMAP:
        case.preconditioner = PC_REMAP[pc]
        case.preconditioner_requested = pc
        if pc != "bjacobi-icc":
            warnings.warn(f"solver.groundwater.preconditioner: {pc} runs as the z-line block "
                          f"Jacobi PCG on this path", stacklevel=2)
    elif pc in ("zline", "jacobi", "chebyshev"):
        case.preconditioner = pc
    else:
        raise ValueError(f"solver.groundwater.preconditioner: unknown value {pc!r}")
    case.precision = str(sg.get("precision", "fp64"))
    # boundary conditions
    entries = []
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dy
    for bc in cfg.get("boundary_conditions", []) or []:
        target = bc["target"]
        if target not in ("groundwater_top", "groundwater_bottom", "groundwater_side"):
            continue
        kind = bc["kind"]
        val = bc.get("value", {"constant": 0.0})
        if kind == "head":
            code = abi.BC_HEAD_HYDROSTATIC if "hydrostatic" in val else abi.BC_HEAD
        elif kind == "flux":
            code = abi.BC_GRAVITY if "gravity" in val else abi.BC_FLUX
        else:
            raise ValueError(f"groundwater boundary condition '{bc['name']}' must be head or flux")
        series = None
        const = 0.0
        if "constant" in val:
            const = float(val["constant"])
        elif "series" in val:
            series = TimeSeries.from_file(base / val["series"]["file"])
        elif "hydrostatic" in val:
            const = float(val["hydrostatic"]["eta"])
        poly = [tuple(map(float, p)) for p in bc["region"]["polygon"]]
        inside = np.array([[polygon_contains(poly, xc[i], yc[j]) for i in range(nx)]
                           for j in range(ny)])
        if target == "groundwater_side":
            for which, sel in (("xm", np.s_[:, 0]), ("xp", np.s_[:, nx - 1]),
                               ("ym", np.s_[0, :]), ("yp", np.s_[ny - 1, :])):
                mask = np.zeros((ny, nx), bool)
                mask[sel] = inside[sel]
                if mask.any():
                    entries.append(BcEntry(which, code, mask, const, series))
        else:
            which = "top" if target == "groundwater_top" else "bot"
            entries.append(BcEntry(which, code, inside, const, series))
    for e in entries:
        case.set_bc(e.which, e.code, e.value(0.0), e.mask)
    return case, BoundarySchedule(entries, ny, nx)


def density_ratios(coupling: dict, salinity=None, temperature=None):
    """Cell ratios from the transport scalars (Baroclinic.cpp:78-86):
    r_rho = 1 + beta_s s - beta_T (T - T0), r_visc = 1 / (1 + beta_sv s);
    an absent scalar contributes 0.  Feed them to RichardsSolver.set_density
    with reallocation_gradient=coupling["reallocation_gradient"]."""
    s = 0.0 if salinity is None else np.asarray(salinity, dtype=np.float64)
    dt = 0.0 if temperature is None else (np.asarray(temperature, dtype=np.float64)
                                          - coupling["reference_temperature"])
    r_rho = 1.0 + coupling["beta_saline"] * s - coupling["thermal_expansion"] * dt
    r_visc = 1.0 / (1.0 + coupling["beta_saline_viscosity"] * s)
    return r_rho, r_visc


def coupled_density_ratios(coupling: dict, salinity=None, temperature=None,
                           surface_salinity=None, surface_temperature=None) -> dict:
    """All four ratio arrays of a coupled run, as keyword arguments of
    RichardsSolver.set_density: the cell ratios from the subsurface scalars
    (halo layout) and the ratios of the ghost above each column's top cell
    from the surface scalars ((ny+2, nx+2) arrays), which the reference takes
    for the top face in coupled runs (Baroclinic.cpp:172-180).  Without the
    top ratios the top face would use the top cell's own ratio, the
    reference's uncoupled rule."""
    r_rho, r_visc = density_ratios(coupling, salinity, temperature)
    top_rho, top_visc = density_ratios(coupling, surface_salinity, surface_temperature)
    shape2 = None
    for a in (surface_salinity, surface_temperature):
        if a is not None:
            shape2 = np.shape(a)
    if shape2 is None:
        raise ValueError("coupled_density_ratios: a surface scalar (salinity or temperature) "
                         "is required for the top-face ratios")
    return dict(r_rho=r_rho, r_visc=r_visc,
                r_rho_top=np.broadcast_to(top_rho, shape2).astype(np.float64),
                r_visc_top=np.broadcast_to(top_visc, shape2).astype(np.float64),
                reallocation_gradient=coupling.get("reallocation_gradient", "freshwater"))


def run_groundwater_loop(solver, schedule: BoundarySchedule, t_end: float, t_start: float = 0.0,
                         on_step=None):
    """The gw-only driver loop (src/driver/Simulation.cpp:1672-1770): t += dtg;
    refresh boundary values at t; step(t, dtg); dtg = nextDt(); mass audit."""
    t = t_start
    dtg = solver.next_dt()
    v_prev = solver.volume()
    records = []
    while t < t_end - 1e-12 * max(1.0, abs(t_end)):
        t += dtg
        if schedule is not None and schedule.time_dependent:
            for which, vals in schedule.values(t).items():
                solver.set_bc_values(which, vals)
        solver.step(dtg, t=t)
        a = solver.audit()
        v = solver.volume()
        a.update(t=t, dtg=dtg, volume=v,
                 residual=(v - v_prev) - (a["boundary_in"] - a["ss_storage"] +
                                          a["realloc_adjust"] - a["vloss"] -
                                          a.get("tr_applied", 0.0)))
        records.append(a)
        if on_step is not None:
            on_step(solver, a)
        v_prev = v
        dtg = solver.next_dt()
    return records
