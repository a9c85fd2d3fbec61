"""Python mirror of frehg::gw::RichardsSolver and frehg::LinearSystem over the
HIP C-ABI (include/frehg_gw.h).

Method names and argument meaning follow the reference
(src/gw/RichardsSolver.hpp:166-423, src/core/LinearSystem.hpp:100-213):
``step(t, dtg)``, ``next_dt()``, ``audit()``, ``volume()`` (ownedVolume),
``last_solve()``, field getters, ``attach_coupling``.  Errors raise
``FrehgError`` where the reference throws ``FatalError``.  Every call goes
to the in-tree libfrehg_gw.so; there is no other code path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import FrehgError, check, load_library

_dp = C.POINTER(C.c_double)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class RichardsSolver:
    """One groundwater domain on the GPU (the reference's RichardsSolver)."""

    def __init__(self, case):
        self.lib = load_library()
        self.case = case
        desc, keep = case.to_desc()
        h = C.c_void_p()
        check(self.lib, self.lib.frehg_gw_create(C.byref(desc), C.byref(h)))
        self._h = h
        self.nx, self.ny, self.nz = case.nx, case.ny, case.nz
        self.t = 0.0
        self._cpl = None
        if case.coupling is not None:
            self.attach_coupling(**case.coupling)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.frehg_gw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def enable_transport_window(self) -> None:
        """RichardsSolver::enableTransportWindow: window volumes from now on."""
        check(self.lib, self.lib.frehg_gw_enable_transport_window(self._h))

    def reset_transport_window(self) -> None:
        """RichardsSolver::resetTransportWindow: zero the window volumes."""
        check(self.lib, self.lib.frehg_gw_reset_transport_window(self._h))

    def field_shape(self, fid):
        if fid in abi.PLANE_FIELDS:
            return (self.ny + 2, self.nx + 2)
        nk = self.nz + 1 if fid in abi.FACE_FIELDS else self.nz
        return (self.ny + 2, self.nx + 2, nk)

    def get(self, fid) -> np.ndarray:
        out = np.empty(self.field_shape(fid), dtype=np.float64)
        check(self.lib, self.lib.frehg_gw_get_field(self._h, fid, _dptr(out)))
        return out

    def set(self, fid, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.field_shape(fid))
        check(self.lib, self.lib.frehg_gw_set_field(self._h, fid, _dptr(a)))

    def set_density(self, r_rho, r_visc=None, r_rho_top=None, r_visc_top=None,
                    reallocation_gradient="freshwater") -> None:
        """Density / viscosity ratios (RichardsSolver::updateBaroclinicFaces,
        Baroclinic.cpp:46-187): r_rho, r_visc per cell in halo layout
        (ny+2, nx+2, nz), ghosts included; r_rho_top / r_visc_top the ghost
        above each column's top cell, (ny+2, nx+2), None for the top cell's
        own.  r_rho None resets every ratio to 1."""
        keep, args = [], []
        shapes = [self.field_shape(abi.F_HEAD)] * 2 + [(self.ny + 2, self.nx + 2)] * 2
        for a, shape in zip((r_rho, r_visc, r_rho_top, r_visc_top), shapes):
            if a is None:
                args.append(None)
            else:
                arr = np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float64)
                keep.append(arr)
                args.append(_dptr(arr))
        check(self.lib, self.lib.frehg_gw_set_density(
            self._h, *args, int(reallocation_gradient == "density")))

    def set_et(self, columns, qvap) -> None:
        """The evapotranspiration soil channel (RichardsSolver::etColumn_ /
        qVap_): a (ny+2, nx+2) mask of channel columns and their applied top
        flux [m/s], positive upward; both None detach it."""
        if (columns is None) != (qvap is None):
            raise FrehgError(abi.FREHG_EINVAL, "set_et: columns and qvap are given together "
                                               "(both None detaches the channel)")
        if columns is None:
            check(self.lib, self.lib.frehg_gw_set_et(self._h, None, None))
            return
        m = np.ascontiguousarray(np.broadcast_to(columns, (self.ny + 2, self.nx + 2)), dtype=np.int32)
        q = np.ascontiguousarray(np.broadcast_to(qvap, (self.ny + 2, self.nx + 2)), dtype=np.float64)
        check(self.lib, self.lib.frehg_gw_set_et(self._h, m.ctypes.data_as(abi._ip), _dptr(q)))

    def set_sink(self, sink) -> None:
        """The root-uptake sink S [1/s] per cell, halo layout (ny+2, nx+2, nz),
        applied from the next step on (RichardsSolver::sink_,
        RichardsSolver.hpp:656); None detaches it."""
        if sink is None:
            check(self.lib, self.lib.frehg_gw_set_field(self._h, abi.F_SINK, None))
        else:
            self.set(abi.F_SINK, sink)

    def set_bc_values(self, which: int, values) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(self.ny, self.nx)
        check(self.lib, self.lib.frehg_gw_set_bc_values(self._h, which, _dptr(a)))

    def attach_coupling(self, eta, dept, seep_accum, avail, gain, bounce=None, min_depth=0.0,
                        vent_all=0):
        """GwCoupling: 2D (ny+2, nx+2) surface arrays, copied to device buffers
        that the step then mutates in place (RichardsSolver.hpp:138-163)."""
        names = ("eta", "dept", "seep_accum", "avail", "gain", "bounce")
        vals = (eta, dept, seep_accum, avail, gain, bounce)
        self._cpl = {n: (abi.DeviceArray(v, self.lib) if v is not None else None)
                     for n, v in zip(names, vals)}
        ptr = lambda n: self._cpl[n].ptr if self._cpl[n] is not None else None
        check(self.lib, self.lib.frehg_gw_attach_coupling(
            self._h, ptr("eta"), ptr("dept"), ptr("seep_accum"), ptr("avail"), ptr("gain"),
            ptr("bounce"), float(min_depth), int(vent_all)))

    def set_coupling_state(self, name, values) -> None:
        """Overwrite one attached coupling array (e.g. after a surface update)."""
        arr = self._cpl[name]
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(arr.shape)
        check(self.lib, self.lib.frehg_gw_synchronize(self._h))
        check(self.lib, self.lib.frehg_memcpy_htod(arr.ptr, a.ctypes.data_as(C.c_void_p),
                                                   arr.nbytes))

    def coupling_state(self, name) -> np.ndarray:
        check(self.lib, self.lib.frehg_gw_synchronize(self._h))
        return self._cpl[name].numpy()

    def step(self, dt=None, t=None) -> float:
        if dt is None:
            dt = self.next_dt()
        self.t = self.t + dt if t is None else t
        check(self.lib, self.lib.frehg_gw_step(self._h, self.t, dt))
        return dt

    def next_dt(self) -> float:
        v = C.c_double()
        check(self.lib, self.lib.frehg_gw_next_dt(self._h, C.byref(v)))
        return v.value

    def audit(self) -> dict:
        a = abi.GwAudit()
        check(self.lib, self.lib.frehg_gw_get_audit(self._h, C.byref(a)))
        return a.as_dict()

    def volume(self) -> float:
        v = C.c_double()
        check(self.lib, self.lib.frehg_gw_owned_volume(self._h, C.byref(v)))
        return v.value

    def last_solve(self) -> dict:
        s = abi.SolveStats()
        check(self.lib, self.lib.frehg_gw_last_solve(self._h, C.byref(s)))
        return {n: getattr(s, n) for n, _ in s._fields_}

    def active_cells(self) -> int:
        v = C.c_int64()
        check(self.lib, self.lib.frehg_gw_active_cells(self._h, C.byref(v)))
        return v.value

    def test_post_allocate(self, dtg: float) -> None:
        """Test export (include/frehg_gw_testing.h): the post-allocation sweep
        and final clamp alone on the current state."""
        check(self.lib, self.lib.frehg_gw_test_post_allocate(self._h, float(dtg)))

    def set_timing(self, enable) -> None:
        """0/False off; n > 0 (True = 1): step and solve events every step and
        one SpMV launch in n bracketed by events."""
        check(self.lib, self.lib.frehg_gw_set_timing(self._h, int(enable)))

    def timing(self) -> dict:
        t = abi.GwTiming()
        check(self.lib, self.lib.frehg_gw_get_timing(self._h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in t._fields_}

    def synchronize(self) -> None:
        check(self.lib, self.lib.frehg_gw_synchronize(self._h))

    def set_step_mode(self, mode, rounds: int = 0, iterations: int = 0) -> None:
        """'host' (host-polled PCG), 'device' (device-controlled, one host
        synchronization a step) or 'graph' (the device-controlled step
        captured in a hipGraph and replayed).  rounds / iterations: the
        device-controlled solve's budget (0, 0: sized from the first step)."""
        modes = {"host": abi.STEP_HOST, "device": abi.STEP_DEVICE, "graph": abi.STEP_GRAPH}
        m = modes[mode] if isinstance(mode, str) else int(mode)
        check(self.lib, self.lib.frehg_gw_set_step_mode(self._h, m, int(rounds), int(iterations)))

    def column_tiles(self) -> list:
        """Depth buckets of the batched column solve: [{R, columns, cpw}]
        (empty when the operator has lateral legs)."""
        cap = 9
        R, n, cpw = (C.c_int * cap)(), (C.c_int * cap)(), (C.c_int * cap)()
        nb = C.c_int()
        check(self.lib, self.lib.frehg_gw_get_column_tiles(self._h, R, n, cpw, cap, C.byref(nb)))
        return [{"R": R[b], "columns": n[b], "cpw": cpw[b]} for b in range(nb.value)]

    def set_chebyshev(self, degree: int = 3, lower_ratio: float = 0.05) -> None:
        check(self.lib, self.lib.frehg_gw_set_chebyshev(self._h, int(degree), float(lower_ratio)))

    def enable_column_solve(self, enable=True) -> None:
        check(self.lib, self.lib.frehg_gw_enable_column_solve(self._h, int(bool(enable))))

    def set_column_tiles(self, cpw) -> None:
        arr = (C.c_int * len(cpw))(*[int(v) for v in cpw])
        check(self.lib, self.lib.frehg_gw_set_column_tiles(self._h, arr, len(cpw)))

    def graph_stats(self) -> dict:
        t = abi.GraphStats()
        check(self.lib, self.lib.frehg_gw_get_graph_stats(self._h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in t._fields_}

    def set_pcg_shapes(self, march64=0, march32=0, column=0) -> None:
        """PCG launch shapes in waves per CU (0 keeps a value); see
        frehg2_amd.autotune."""
        check(self.lib, self.lib.frehg_gw_set_pcg_shapes(self._h, int(march64), int(march32),
                                                         int(column)))

    def pcg_shapes(self) -> dict:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(self.lib, self.lib.frehg_gw_get_pcg_shapes(self._h, C.byref(a), C.byref(b),
                                                         C.byref(c)))
        return {"march64": a.value, "march32": b.value, "column": c.value}


class LinearSystem:
    """The COO solve seam (LinearSystem::setPattern / setValues / solve).
    Values, rhs and solution are device arrays (abi.DeviceArray)."""

    def __init__(self, n_rows, rtol=1e-8, atol=1e-14, max_iterations=500, pc=abi.PC_JACOBI):
        self.lib = load_library()
        h = C.c_void_p()
        check(self.lib, self.lib.frehg_ls_create(n_rows, rtol, atol, max_iterations, pc,
                                                 C.byref(h)))
        self._h = h
        self.n = n_rows

    def __del__(self):
        try:
            self.lib.frehg_ls_destroy(self._h)
        except Exception:
            pass

    def set_pattern(self, rows, cols):
        r = np.ascontiguousarray(rows, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        check(self.lib, self.lib.frehg_ls_set_pattern(
            self._h, r.size, r.ctypes.data_as(C.POINTER(C.c_int64)),
            c.ctypes.data_as(C.POINTER(C.c_int64))))

    def set_values(self, values_device):
        check(self.lib, self.lib.frehg_ls_set_values(self._h, values_device.ptr))

    def solve(self, rhs_device, sol_device) -> dict:
        st = abi.SolveStats()
        check(self.lib, self.lib.frehg_ls_solve(self._h, rhs_device.ptr, sol_device.ptr,
                                                C.byref(st)))
        return {n: getattr(st, n) for n, _ in st._fields_}

    def is_symmetric(self, tol) -> bool:
        s = C.c_int()
        check(self.lib, self.lib.frehg_ls_is_symmetric(self._h, tol, C.byref(s)))
        return bool(s.value)

    def set_precision(self, precision, inner_rtol=0.0) -> None:
        """FP64 (default) or the FP32-inner mixed form (structured path only)."""
        prec = {"fp64": abi.PREC_FP64, "mixed": abi.PREC_MIXED}.get(precision, precision)
        check(self.lib, self.lib.frehg_ls_set_precision(self._h, int(prec), float(inner_rtol)))

    def path(self) -> dict:
        """Which path runs: 'structured' (the 7-point stencil PCG), 'pattern'
        (the 7-point pattern, values not set or not symmetric) or 'generic'."""
        s, a = C.c_int(), C.c_double()
        check(self.lib, self.lib.frehg_ls_structured(self._h, C.byref(s), C.byref(a)))
        return {"path": ("generic", "pattern", "structured")[s.value], "asymmetry": a.value}
