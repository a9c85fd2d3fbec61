"""Shared fixtures: the reference's MiniGw configurations as GwCase objects.

gw_base_case mirrors gwBaseConfig (reference tests/unit/test_gw.cpp:86-109):
dx = dy = 1, flat bed at 0, fixed dtg, Ss = 0, the soft test soil
(testSoil, :71-84) and a uniform moisture IC of 0.2.
"""
import numpy as np

from frehg2_amd import abi
from frehg2_amd.cases import GwCase

TEST_SOIL = dict(theta_s=0.46, theta_r=0.04, vg_alpha=5.9, vg_n=2.68, aev=-0.02)


def gw_base_case(nx, ny, nz, dz, dtg, ks=1.0e-5, **kw):
    c = GwCase(name="mini", nx=nx, ny=ny, nz=nz, dx=1.0, dy=1.0, dz=dz,
               ksx=ks, ksy=ks, ksz=ks, specific_storage=0.0,
               dt_init=dtg, dt_min=dtg, dt_max=dtg, ic_form="moisture", ic_const=0.2,
               rtol=1e-8, atol=1e-14, max_iterations=1000, dt=dtg, **TEST_SOIL)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def interior(field):
    return field[1:-1, 1:-1, :]


def run_steps(sim, n, fn=None):
    """Adaptive steps as MiniGw::step (dtg = nextDt()); fn(sim, dtg) per step."""
    total = 0.0
    for s in range(n):
        if getattr(sim.case, "bc_schedule", None) is not None:
            for which, vals in sim.case.bc_schedule(s).items():
                sim.set_bc_values(which, vals)
        dtg = sim.step()
        total += dtg
        if fn is not None:
            fn(sim, dtg)
    return total
