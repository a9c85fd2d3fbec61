"""The allocation-free solve (frehg_ls_solve_stencil_noalloc) against the
solve seam's structured path (frehg_ls_solve) on the same systems.

Both run the same FP64 stencil PCG kernels at the same launch shapes, the
seam with its own allocated workspace and the allocation-free entry with the
library's static scratch array, so the solutions and iteration counts agree
bitwise.  Two calls in a row, the second on a smaller grid (its ghost cells
sit where the first call's interior values were), then a grid beyond the
scratch maximum, which is refused before any device work.  Systems are the
oracle's assembly of a terrain-following slice of the 26.2 M-cell workload
(every column full, so the seam's embedding is the grid itself).
"""
import ctypes as C

import numpy as np
import pytest

from frehg2_amd import abi, cases
from frehg2_amd.abi import DeviceArray

pytestmark = pytest.mark.gpu


def _system(nx, ny, nz):
    from oracle.oracle import OracleRichards
    case = cases.hetero_512x512x100(nx=nx, ny=ny, nz=nz)
    vals, rhs, cols = OracleRichards(case).assemble_only(300.0)
    return vals, rhs, cols


def _stencil(nx, ny, nz, vals, rhs, cols):
    """The COO system in halo-layout stencil storage (rows in (j, i, k) order)."""
    shape = (ny + 2, nx + 2, nz)
    Ad = np.ones(shape)
    Axp, Ayp, Azp, b = (np.zeros(shape) for _ in range(4))
    inner = (slice(1, -1), slice(1, -1), slice(None))
    v = vals.reshape(ny, nx, nz, 7)
    c = cols.reshape(ny, nx, nz, 7)
    Ad[inner] = v[..., 0]
    Axp[inner] = np.where(c[..., 2] >= 0, v[..., 2], 0.0)
    Ayp[inner] = np.where(c[..., 4] >= 0, v[..., 4], 0.0)
    Azp[inner] = np.where(c[..., 6] >= 0, v[..., 6], 0.0)
    b[inner] = rhs.reshape(ny, nx, nz)
    return Ad, Axp, Ayp, Azp, b


def _noalloc(nx, ny, nz, arrays, pc):
    lib = abi.load_library()
    dev = [DeviceArray(a) for a in arrays]
    sol = DeviceArray.zeros(arrays[0].shape)
    st = abi.SolveStats()
    print(f"[noalloc] grid {nx}x{ny}x{nz} pc={pc}", flush=True)
    rc = lib.frehg_ls_solve_stencil_noalloc(nx, ny, nz, dev[0].ptr, dev[1].ptr, dev[2].ptr,
                                            dev[3].ptr, dev[4].ptr, sol.ptr, 1e-12, 1e-300, 5000,
                                            pc, C.byref(st))
    abi.check(lib, rc)
    return sol.numpy()[1:-1, 1:-1].reshape(-1), st


def _seam(n, vals, rhs, cols, pc):
    from frehg2_amd.gw import LinearSystem
    ls = LinearSystem(n, rtol=1e-12, atol=1e-300, max_iterations=5000, pc=pc)
    ls.set_pattern(np.repeat(np.arange(n), 7), cols.reshape(-1))
    ls.set_values(DeviceArray(vals.reshape(-1)))
    assert ls.path()["path"] == "structured"
    xd = DeviceArray.zeros(n)
    st = ls.solve(DeviceArray(rhs), xd)
    return xd.numpy(), st


@pytest.mark.parametrize("pc", [abi.PC_ZLINE, abi.PC_JACOBI])
def test_noalloc_solve_matches_the_seam_on_two_calls(pc):
    for nx, ny, nz in ((40, 30, 100), (24, 18, 60)):  # the second call smaller
        vals, rhs, cols = _system(nx, ny, nz)
        n = rhs.size
        x_seam, st_seam = _seam(n, vals, rhs, cols, pc)
        x_na, st_na = _noalloc(nx, ny, nz, _stencil(nx, ny, nz, vals, rhs, cols), pc)
        print(f"[noalloc] {nx}x{ny}x{nz}: {st_na.iterations} it (seam {st_seam['iterations']})",
              flush=True)
        assert st_na.converged == 1
        assert st_na.iterations == st_seam["iterations"]
        assert np.array_equal(x_na, x_seam)


def test_noalloc_refuses_grids_beyond_its_scratch():
    lib = abi.load_library()
    st = abi.SolveStats()
    small = DeviceArray.zeros(8)  # the size check comes before any device work
    p = small.ptr
    rc = lib.frehg_ls_solve_stencil_noalloc(200, 200, 100, p, p, p, p, p, p, 1e-12, 0.0, 10,
                                            abi.PC_ZLINE, C.byref(st))
    assert rc == abi.FREHG_EINVAL
    assert b"scratch" in lib.frehg_last_error()
    assert (200 + 2) * (200 + 2) * 100 > abi.NOALLOC_MAX_CELLS


def test_noalloc_misaligned_operands_take_the_one_row_kernels():
    """Every operand one double into its buffer (8-byte, not 16-byte
    aligned): the FP64 pair kernels load and store 16-byte vectors, so the
    solve runs the one-row-a-lane kernels instead.  Same iterations and the
    aligned solve's solution to rounding; bitwise equality would mean the
    pair kernels ran on the misaligned operands."""
    nx, ny, nz = 40, 30, 100
    vals, rhs, cols = _system(nx, ny, nz)
    arrays = _stencil(nx, ny, nz, vals, rhs, cols)
    x_al, st_al = _noalloc(nx, ny, nz, arrays, abi.PC_ZLINE)
    lib = abi.load_library()
    dev = [DeviceArray(np.concatenate([[0.0], a.reshape(-1)])) for a in arrays]
    sol = DeviceArray.zeros(arrays[0].size + 1)
    st = abi.SolveStats()
    print(f"[noalloc] grid {nx}x{ny}x{nz} misaligned", flush=True)
    p = [d.ptr.value + 8 for d in dev]
    rc = lib.frehg_ls_solve_stencil_noalloc(nx, ny, nz, p[0], p[1], p[2], p[3], p[4],
                                            sol.ptr.value + 8, 1e-12, 1e-300, 5000,
                                            abi.PC_ZLINE, C.byref(st))
    abi.check(lib, rc)
    x_mis = sol.numpy()[1:].reshape(arrays[0].shape)[1:-1, 1:-1].reshape(-1)
    err = float(np.max(np.abs(x_mis - x_al)) / np.max(np.abs(x_al)))
    print(f"[noalloc] misaligned {st.iterations} it (aligned {st_al.iterations}), "
          f"difference {err:.2e}", flush=True)
    assert st.converged == 1 and abs(st.iterations - st_al.iterations) <= 1
    assert err <= 1e-10
    assert not np.array_equal(x_mis, x_al)
