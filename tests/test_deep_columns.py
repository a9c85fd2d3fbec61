"""Columns deeper than 512 cells (513 .. 1024 rows run the column kernels'
deep form, R = 16 row segments of 64, rows past the column masked).

The reference has no depth limit; here nz <= 1024.  Identical-input steps
against the FP64 oracle at the parity bar (head, theta, fluxes 1e-10; the
mixed form 1e-8), on a laterally coupled grid (PCG, z-line and point
Jacobi) and on uneven decoupled columns reaching 1000 cells (the direct
column solve's deep bucket); and an infiltrating 1000-cell column whose
audit closes.  The validation (nz > 1024 rejected) runs on the CPU.
"""
import numpy as np
import pytest

from frehg2_amd import abi, cases
from helpers import gw_base_case


def test_depth_limit_is_validated():
    """Checked before any device work, so it runs without a GPU."""
    from frehg2_amd.gw import RichardsSolver
    c = gw_base_case(1, 1, 1025, 0.01, 10.0)
    with pytest.raises(abi.FrehgError) as e:
        RichardsSolver(c)
    assert e.value.code == abi.FREHG_EUNSUPPORTED and "1024" in str(e.value)


def _deep(nx, ny, nz, precision="fp64", pc="zline"):
    c = gw_base_case(nx, ny, nz, 0.01, 30.0, ic_const=0.25)
    c.name = f"deep_{nx}x{ny}x{nz}"
    c.precision = precision
    c.preconditioner = pc
    c.set_bc("top", abi.BC_FLUX, -1.0e-5)
    c.set_bc("bot", abi.BC_GRAVITY, 0.0)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("nz,precision,pc", [(700, "fp64", "zline"), (1000, "fp64", "jacobi"),
                                             (1000, "mixed", "zline"), (770, "mixed", "zline")])
def test_deep_grid_matches_the_oracle(nz, precision, pc):
    from test_gpu_parity import _stepwise
    case = _deep(4, 3, nz, precision, pc)
    tol = 1e-10 if precision == "fp64" else 1e-8
    o, g = _stepwise(case, 4, tol=tol)
    assert g.active_cells() == o.active_cells() == 4 * 3 * nz


@pytest.mark.gpu
def test_deep_uneven_columns_take_the_deep_bucket():
    """Decoupled columns of 300 .. 1000 cells: the column solve's depth
    buckets include the deep one (R = 16), and each step matches the
    oracle's."""
    from test_gpu_parity import _stepwise
    from frehg2_amd.gw import RichardsSolver
    c = cases.columns_4096(ncol_side=6, lmax=1000)
    rng = np.random.Generator(np.random.Philox(11))
    depth = rng.integers(300, 1001, (6, 6))
    depth[0, 0] = 1000
    c.depth = depth
    c.bath = -(1000 - depth) * 0.01
    g = RichardsSolver(c)
    Rs = [t["R"] for t in g.column_tiles()]
    assert Rs[0] == 16 and sum(t["columns"] for t in g.column_tiles()) == 36
    g.close()
    _stepwise(c, 4)


@pytest.mark.gpu
def test_deep_infiltrating_column_closes_its_audit():
    from frehg2_amd.gw import RichardsSolver
    c = _deep(1, 1, 1000)
    g = RichardsSolver(c)
    for _ in range(20):
        v0 = g.volume()
        g.step(30.0)
        a = g.audit()
        res = (g.volume() - v0) - (a["boundary_in"] - a["ss_storage"] + a["realloc_adjust"]
                                   - a["vloss"])
        assert abs(res) <= 1e-12 * g.volume(), res
    assert a["boundary_in"] > 0.0
