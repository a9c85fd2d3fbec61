"""Groundwater cases: the descriptor builder and the seeded synthetic workloads.

A ``GwCase`` holds everything the reference's RichardsSolver constructor reads
(src/gw/RichardsSolver.cpp:61-205) as numpy arrays and converts it to the
``frehg_gw_desc`` of include/frehg_gw.h.  The five workloads are the named
configurations of this path:

  richards_loam_32x32x50        homogeneous loam column block, ponded top
  richards_hetero_512x512x100   26.2 M cells, terrain-following, log-normal Ks
  richards_stiff_32x32x50       stiff contrasts for the mixed-precision solve
  richards_columns_4096         4096 uneven 1D columns (lateral Ks = 0)
  richards_coupled_128x128x64   one ensemble member with a ponded surface

Seeded fields use numpy's counter-based Philox generator with fixed seeds
(20261203 + workload index), so every host regenerates bit-identical inputs.
Smooth random fields are white noise filtered by a Gaussian in Fourier space
(periodic), renormalised to zero mean and unit variance; a correlation
length L cells means a filter standard deviation of L/2 cells, i.e. the
filtered field's autocorrelation exp(-r^2 / L^2) falls to 1/e at r = L.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi

SEED_BASE = 20261203


def _rng(index: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(SEED_BASE + index))


def smooth_field(rng: np.random.Generator, shape, corr_cells: float) -> np.ndarray:
    """Zero-mean, unit-variance Gaussian random field with correlation length
    ``corr_cells`` (Gaussian filter sigma = corr_cells / 2), periodic."""
    white = rng.standard_normal(shape)
    sigma = corr_cells / 2.0
    spec = np.fft.rfftn(white)
    freqs = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    damp = np.ones(spec.shape)
    for ax, f in enumerate(freqs):
        g = np.exp(-2.0 * (np.pi * sigma * f) ** 2)
        sh = [1] * len(shape)
        sh[ax] = f.size
        damp = damp * g.reshape(sh)
    out = np.fft.irfftn(spec * damp, s=shape, axes=list(range(len(shape))))
    out -= out.mean()
    std = out.std()
    return out / std if std > 0 else out


@dataclass
class GwCase:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dz_stretch: float = 1.0
    follow_terrain: bool = False
    terrain_layers: str = "scaled"
    bath: Optional[np.ndarray] = None          # (ny, nx)
    bath_const: float = 0.0
    # soil: scalars or arrays broadcastable to (ny, nx, nz)
    ksx: object = 1e-5
    ksy: object = 1e-5
    ksz: object = 1e-5
    vg_alpha: object = 5.9
    vg_n: object = 2.68
    theta_s: object = 0.46
    theta_r: object = 0.04
    aev: object = 0.0
    specific_storage: float = 0.0
    use_full3d: bool = True
    face_conductivity: str = "arithmetic"
    predictor_corrector: str = "consistent"
    density_coupling: Optional[dict] = None  # groundwater.density_coupling (betas, gradient)
    reallocation: str = "legacy"
    reallocation_surplus: str = "drop"
    dt_init: float = 1.0
    dt_min: float = 1.0
    dt_max: float = 1.0
    dq_grow: float = 0.01
    dq_shrink: float = 0.02
    courant_max: float = 2.0
    ic_form: str = "moisture"
    ic_values: Optional[np.ndarray] = None     # (ny,nx,nz) or (ny,nx) for water_table
    ic_const: float = 0.2
    bc: dict = field(default_factory=dict)     # name -> (codes (ny,nx), values (ny,nx))
    rtol: float = 1e-8
    atol: float = 1e-14
    max_iterations: int = 1000
    preconditioner: str = "zline"
    precision: str = "fp64"
    initial_guess_previous: bool = True
    inner_rtol: float = 0.0
    # workload metadata (not part of the descriptor)
    steps: int = 1
    dt: float = 1.0
    coupling: Optional[dict] = None            # 2D (ny+2, nx+2) arrays + min_depth
    bc_schedule: Optional[object] = None       # callable(step) -> {which: values}
    adaptive: bool = False                     # steps follow next_dt() (else dt)

    def soil_array(self, name: str) -> np.ndarray:
        v = getattr(self, name)
        a = np.broadcast_to(np.asarray(v, dtype=np.float64), (self.ny, self.nx, self.nz))
        return np.ascontiguousarray(a, dtype=np.float64)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def set_bc(self, which: str, code: int, value=0.0, mask: Optional[np.ndarray] = None):
        """Mark a boundary (top/bottom/xm/xp/ym/yp) with one code and value(s)."""
        codes = np.zeros((self.ny, self.nx), dtype=np.int32)
        vals = np.zeros((self.ny, self.nx), dtype=np.float64)
        if which in self.bc:
            codes[...] = self.bc[which][0]
            vals[...] = self.bc[which][1]
        sel = np.ones((self.ny, self.nx), dtype=bool) if mask is None else mask
        codes[sel] = code
        vals[sel] = np.broadcast_to(np.asarray(value, dtype=np.float64), (self.ny, self.nx))[sel]
        self.bc[which] = (codes, vals)

    def to_desc(self):
        """Return (GwDesc, keepalive) for the C-ABI; keepalive must outlive the call."""
        keep = []

        def dptr(a):
            arr = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(C.POINTER(C.c_double))

        def iptr(a):
            arr = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(arr)
            return arr.ctypes.data_as(C.POINTER(C.c_int))

        d = abi.GwDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.dx, d.dy, d.dz = self.dx, self.dy, self.dz
        d.dz_stretch = self.dz_stretch
        d.follow_terrain = int(self.follow_terrain)
        d.terrain_layers = 1 if self.terrain_layers == "uniform" else 0
        if self.bath is not None:
            d.bottom_elevation = dptr(self.bath)
        d.bottom_elevation_const = self.bath_const
        for n in ("ksx", "ksy", "ksz", "vg_alpha", "vg_n", "theta_s", "theta_r", "aev"):
            setattr(d, n, dptr(self.soil_array(n)))
        d.specific_storage = self.specific_storage
        d.use_full3d = int(self.use_full3d)
        d.predictor_corrector = 0 if self.predictor_corrector == "legacy" else 1
        d.face_conductivity = 1 if self.face_conductivity == "harmonic" else 0
        d.saturated_contact = 1
        d.reallocation = 1 if self.reallocation == "disabled" else 0
        d.reallocation_surplus = 1 if self.reallocation_surplus == "redistribute" else 0
        d.dt_init, d.dt_min, d.dt_max = self.dt_init, self.dt_min, self.dt_max
        d.dq_grow, d.dq_shrink, d.courant_max = self.dq_grow, self.dq_shrink, self.courant_max
        d.ic_form = {"moisture": abi.IC_MOISTURE, "head": abi.IC_HEAD,
                     "water_table": abi.IC_WATER_TABLE}[self.ic_form]
        if self.ic_values is not None:
            d.ic_values = dptr(self.ic_values)
        d.ic_const = self.ic_const
        for which in ("top", "bot", "xm", "xp", "ym", "yp"):
            if which in self.bc:
                codes, vals = self.bc[which]
                setattr(d, f"{which}_code", iptr(codes))
                setattr(d, f"{which}_value", dptr(vals))
        d.rtol, d.atol, d.max_iterations = self.rtol, self.atol, self.max_iterations
        d.preconditioner = {"jacobi": abi.PC_JACOBI, "zline": abi.PC_ZLINE,
                            "icc0": abi.PC_ICC0,
                            "chebyshev": abi.PC_CHEBYSHEV}[self.preconditioner]
        d.precision = abi.PREC_MIXED if self.precision == "mixed" else abi.PREC_FP64
        d.initial_guess_previous = int(self.initial_guess_previous)
        d.inner_rtol = self.inner_rtol
        return d, keep


BC_CODES = {"noflux": abi.BC_NOFLUX, "head": abi.BC_HEAD, "flux": abi.BC_FLUX,
            "gravity": abi.BC_GRAVITY, "hydrostatic": abi.BC_HEAD_HYDROSTATIC}

LOAM = dict(theta_s=0.43, theta_r=0.078, vg_alpha=3.6, vg_n=1.56, aev=0.0)


def loam_32x32x50(steps: int = 100, nx: int = 32, ny: int = 32, nz: int = 50) -> GwCase:
    """Plain FP64 homogeneous loam: 32x32x50, dx=dy=1, dz=0.02, h0=-2, ponded
    head 0.01 on top, free drainage at the bottom, no-flux sides, dt=60."""
    c = GwCase(name="richards_loam_32x32x50", nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=0.02,
               ksx=2.89e-6, ksy=2.89e-6, ksz=2.89e-6, specific_storage=1e-5,
               dt_init=60.0, dt_min=60.0, dt_max=60.0, ic_form="head", ic_const=-2.0,
               rtol=1e-12, atol=1e-30, steps=steps, dt=60.0, **LOAM)
    c.set_bc("top", abi.BC_HEAD, 0.01)
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)
    return c


def hetero_512x512x100(steps: int = 50, nx: int = 512, ny: int = 512, nz: int = 100,
                       precision: str = "fp64") -> GwCase:
    """26.2 M-cell terrain-following case: dx=dy=2, uniform dz=0.05 below a 1 %
    slope bed with a seeded smooth perturbation (0.5 m std, 20-cell length),
    log-normal Ks (ln-mean ln 5e-6, sigma 1.0, 20-cell correlation), alpha
    U[1,8], n U[1.2,2.5]; hydrostatic IC from a table 2 m below the bed; top
    rain 1e-6 m/s (flux, negative = downward); bottom head 3.0 m; 50 x 300 s."""
    rng = _rng(2)
    x = (np.arange(nx) + 0.5) * 2.0
    bump = smooth_field(rng, (ny, nx), 20.0)
    bath = 0.01 * x[None, :] + 0.5 * bump
    lnk = np.log(5e-6) + 1.0 * smooth_field(rng, (ny, nx, nz), 20.0)
    ks = np.exp(lnk)
    alpha = rng.uniform(1.0, 8.0, (ny, nx, nz))
    n = rng.uniform(1.2, 2.5, (ny, nx, nz))
    c = GwCase(name="richards_hetero_512x512x100", nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=0.05,
               follow_terrain=True, terrain_layers="uniform", bath=bath,
               ksx=ks, ksy=ks, ksz=ks, vg_alpha=alpha, vg_n=n, theta_s=0.43, theta_r=0.078,
               aev=0.0, specific_storage=1e-5, dt_init=300.0, dt_min=300.0, dt_max=300.0,
               ic_form="water_table", ic_values=bath - 2.0, rtol=1e-10, atol=1e-30,
               steps=steps, dt=300.0, precision=precision)
    c.set_bc("top", abi.BC_FLUX, -1e-6)
    c.set_bc("bot", abi.BC_HEAD, 3.0)
    return c


def stiff_32x32x50(steps: int = 100, precision: str = "mixed") -> GwCase:
    """The loam grid with stiffer contrasts: sigma_lnK 2.0 (8-cell correlation),
    n U[1.1,3.0], alpha U[1,8]; uniform h0 = -5 (the dry profile above a table
    that lies below this 1 m column); ponded head 0.05 on top; free drainage;
    IR to rtol 1e-10.  At a fixed 60 s step the PCA scheme diverges on these
    contrasts (the FP64 oracle does too), so the step follows the adaptive
    controller (AdaptiveStep.cpp) from 1 s within [0.01, 60] s."""
    rng = _rng(3)
    nx = ny = 32
    nz = 50
    ks = np.exp(np.log(2.89e-6) + 2.0 * smooth_field(rng, (ny, nx, nz), 8.0))
    alpha = rng.uniform(1.0, 8.0, (ny, nx, nz))
    n = rng.uniform(1.1, 3.0, (ny, nx, nz))
    c = GwCase(name="richards_stiff_32x32x50", nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=0.02,
               ksx=ks, ksy=ks, ksz=ks, vg_alpha=alpha, vg_n=n, theta_s=0.43, theta_r=0.078,
               aev=0.0, specific_storage=1e-5, dt_init=1.0, dt_min=0.01, dt_max=60.0,
               ic_form="head", ic_const=-5.0, rtol=1e-10, atol=1e-30, steps=steps, dt=1.0,
               precision=precision)
    c.adaptive = True
    c.set_bc("top", abi.BC_HEAD, 0.05)
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)
    return c


def columns_4096(steps: int = 50, ncol_side: int = 64, lmax: int = 400) -> GwCase:
    """4096 independent soil columns of depth U{50..400} cells (dz=0.01) as a
    64x64x400 regular grid with lateral Ks = 0 and each bed lowered by
    (400-L) dz (ktop masking); per-column van Genuchten parameters, initial
    head U[-10,-0.5], rain pulses of 0-2e-5 m/s, free drainage; adaptive step
    from 1 s within [0.01, 60] s (a fixed 60 s step diverges at dz = 1 cm)."""
    rng = _rng(4)
    ny = nx = ncol_side
    nz = lmax
    depth = rng.integers(50, lmax + 1, (ny, nx))
    bath = -(lmax - depth) * 0.01
    shp = (ny, nx, 1)
    ths = rng.uniform(0.35, 0.5, shp)
    thr = rng.uniform(0.02, 0.1, shp)
    alpha = np.exp(rng.uniform(np.log(0.5), np.log(15.0), shp))
    n = rng.uniform(1.1, 2.8, shp)
    ks = np.exp(rng.uniform(np.log(1e-7), np.log(1e-4), shp))
    h0 = rng.uniform(-10.0, -0.5, shp)
    amp = rng.uniform(0.0, 2e-5, (ny, nx))
    period = rng.integers(5, 21, (ny, nx))
    phase = rng.integers(0, 20, (ny, nx)) % period
    c = GwCase(name="richards_columns_4096", nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=0.01,
               bath=bath, ksx=0.0, ksy=0.0, ksz=np.broadcast_to(ks, (ny, nx, nz)),
               vg_alpha=np.broadcast_to(alpha, (ny, nx, nz)),
               vg_n=np.broadcast_to(n, (ny, nx, nz)),
               theta_s=np.broadcast_to(ths, (ny, nx, nz)),
               theta_r=np.broadcast_to(thr, (ny, nx, nz)), aev=0.0, specific_storage=1e-5,
               dt_init=1.0, dt_min=0.01, dt_max=60.0, ic_form="head",
               ic_values=np.broadcast_to(h0, (ny, nx, nz)).copy(), rtol=1e-10, atol=1e-30,
               steps=steps, dt=1.0)
    c.adaptive = True
    c.set_bc("top", abi.BC_FLUX, 0.0)
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)

    def schedule(step: int):
        on = ((step + phase) // period) % 2 == 0
        return {0: np.where(on, -amp, 0.0)}

    c.bc_schedule = schedule
    c.depth = depth
    return c


def coupled_128x128x64(member: int = 0, steps: int = 20, nx: int = 128, ny: int = 128,
                       nz: int = 64, precision: str = "mixed") -> GwCase:
    """One member of the coupled ensemble: 128x128x64, dx=dy=1, dz=0.05,
    log-normal Ks (ln-mean ln 5e-6, sigma 1.5, 10-cell correlation), alpha
    U[1,8], n U[1.2,2.5]; water table 1 m below a flat surface; a ponded-depth
    field of 0-0.1 m on about half of the columns feeds the coupled top
    (dept, avail = dept, eta = dept), seepage faces elsewhere; adaptive step
    from 1 s within [0.01, 60] s."""
    rng = _rng(5 + 1000 * member)
    ks = np.exp(np.log(5e-6) + 1.5 * smooth_field(rng, (ny, nx, nz), 10.0))
    alpha = rng.uniform(1.0, 8.0, (ny, nx, nz))
    n = rng.uniform(1.2, 2.5, (ny, nx, nz))
    pond = np.clip(0.1 * smooth_field(rng, (ny, nx), 16.0) * 0.5, 0.0, 0.1)
    c = GwCase(name="richards_coupled_128x128x64", nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=0.05,
               ksx=ks, ksy=ks, ksz=ks, vg_alpha=alpha, vg_n=n, theta_s=0.43, theta_r=0.078,
               aev=0.0, specific_storage=1e-5, dt_init=1.0, dt_min=0.01, dt_max=60.0,
               ic_form="water_table", ic_const=-1.0, rtol=1e-10, atol=1e-30, steps=steps,
               dt=1.0, precision=precision)
    c.adaptive = True
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)
    dept = np.zeros((ny + 2, nx + 2))
    dept[1:-1, 1:-1] = pond
    c.coupling = dict(eta=dept.copy(), dept=dept.copy(), seep_accum=np.zeros_like(dept),
                      avail=dept.copy(), gain=np.zeros_like(dept), bounce=np.zeros_like(dept),
                      min_depth=1e-4, vent_all=0)
    return c


WORKLOADS = {
    "richards_loam_32x32x50": loam_32x32x50,
    "richards_hetero_512x512x100": hetero_512x512x100,
    "richards_stiff_32x32x50": stiff_32x32x50,
    "richards_columns_4096": columns_4096,
    "richards_coupled_128x128x64": coupled_128x128x64,
}
