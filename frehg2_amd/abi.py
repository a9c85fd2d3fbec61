"""ctypes mirror of include/frehg_gw.h (the C-ABI of the MI355X groundwater step).

The structures here are byte-for-byte the C structs of the public header; the
loader finds the in-tree HIP library and fails loudly when it is missing, so no
code path can silently fall back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libfrehg_gw.so"

ABI_VERSION = 1

FREHG_OK = 0
FREHG_EINVAL = 1
FREHG_EDEVICE = 2
FREHG_EDIVERGED = 3
FREHG_EUNSUPPORTED = 4

BC_NOFLUX, BC_HEAD, BC_FLUX, BC_GRAVITY, BC_HEAD_HYDROSTATIC = 0, 1, 2, 3, 4
IC_MOISTURE, IC_HEAD, IC_WATER_TABLE = 0, 1, 2
PC_JACOBI, PC_ZLINE, PC_ICC0, PC_CHEBYSHEV = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
STEP_HOST, STEP_DEVICE, STEP_GRAPH = 0, 1, 2

F_HEAD, F_WC, F_HEAD_N, F_WC_N = 0, 1, 2, 3
F_KX, F_KY, F_KZF = 4, 5, 6
F_QX, F_QY, F_QZF = 7, 8, 9
F_VLOSS, F_DZ3D, F_BOT3D = 10, 11, 12
F_AREA_X, F_AREA_Y, F_SIN_X, F_COS_X, F_SIN_Y, F_COS_Y = 13, 14, 15, 16, 17, 18
F_CAPACITY = 19
F_SINK = 20
F_WIN_X, F_WIN_Y, F_WIN_Z, F_WIN_SINK, F_QZ_PATH_CORR, F_WIN_VENT, F_WIN_VAP = range(21, 28)

FACE_FIELDS = (F_KZF, F_QZF, F_WIN_Z, F_QZ_PATH_CORR)
PLANE_FIELDS = (F_WIN_VENT, F_WIN_VAP)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class GwDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dz_stretch", C.c_double),
        ("follow_terrain", C.c_int), ("terrain_layers", C.c_int),
        ("bottom_elevation", _dp), ("bottom_elevation_const", C.c_double),
        ("ksx", _dp), ("ksy", _dp), ("ksz", _dp), ("vg_alpha", _dp), ("vg_n", _dp),
        ("theta_s", _dp), ("theta_r", _dp), ("aev", _dp),
        ("specific_storage", C.c_double),
        ("use_full3d", C.c_int), ("predictor_corrector", C.c_int),
        ("face_conductivity", C.c_int), ("saturated_contact", C.c_int),
        ("reallocation", C.c_int), ("reallocation_surplus", C.c_int),
        ("dt_init", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("dq_grow", C.c_double), ("dq_shrink", C.c_double), ("courant_max", C.c_double),
        ("ic_form", C.c_int), ("ic_values", _dp), ("ic_const", C.c_double),
        ("top_code", _ip), ("bot_code", _ip), ("xm_code", _ip), ("xp_code", _ip),
        ("ym_code", _ip), ("yp_code", _ip),
        ("top_value", _dp), ("bot_value", _dp), ("xm_value", _dp), ("xp_value", _dp),
        ("ym_value", _dp), ("yp_value", _dp),
        ("rtol", C.c_double), ("atol", C.c_double), ("max_iterations", C.c_int),
        ("preconditioner", C.c_int), ("precision", C.c_int),
        ("initial_guess_previous", C.c_int), ("inner_rtol", C.c_double),
    ]


class GwAudit(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "boundary_in", "ss_storage", "realloc_adjust", "realloc_dropped", "vloss",
        "cpl_exchanged", "cpl_bounce", "cpl_vent", "cpl_evap", "cpl_vent_thin", "pc_clamp",
        "tr_applied")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("refinements", C.c_int), ("converged", C.c_int),
                ("residual_norm", C.c_double), ("rhs_norm", C.c_double),
                ("fp64_iterations", C.c_int)]


class GwTiming(C.Structure):
    _fields_ = [("step_ms", C.c_double), ("solve_ms", C.c_double), ("spmv_ms", C.c_double),
                ("spmv_launches", C.c_int64), ("spmv_bytes", C.c_double),
                ("steps", C.c_int64), ("pcg_iterations", C.c_int64)]


class GraphStats(C.Structure):
    _fields_ = [("mode", C.c_int), ("budget_rounds", C.c_int), ("budget_iterations", C.c_int),
                ("captures", C.c_int64), ("replays", C.c_int64), ("device_steps", C.c_int64),
                ("host_finishes", C.c_int64), ("budget_later", C.c_int)]


# Exported symbols of the product library, with argtypes (the C-ABI surface).
_vp = C.c_void_p
SYMBOLS = {
    "frehg_abi_version": ([], C.c_int),
    "frehg_last_error": ([], C.c_char_p),
    "frehg_gw_create": ([C.POINTER(GwDesc), C.POINTER(_vp)], C.c_int),
    "frehg_gw_destroy": ([_vp], C.c_int),
    "frehg_gw_step": ([_vp, C.c_double, C.c_double], C.c_int),
    "frehg_gw_next_dt": ([_vp, _dp], C.c_int),
    "frehg_gw_get_audit": ([_vp, C.POINTER(GwAudit)], C.c_int),
    "frehg_gw_owned_volume": ([_vp, _dp], C.c_int),
    "frehg_gw_last_solve": ([_vp, C.POINTER(SolveStats)], C.c_int),
    "frehg_gw_active_cells": ([_vp, C.POINTER(C.c_int64)], C.c_int),
    "frehg_gw_field_size": ([_vp, C.c_int, C.POINTER(C.c_int64)], C.c_int),
    "frehg_gw_get_field": ([_vp, C.c_int, _dp], C.c_int),
    "frehg_gw_set_field": ([_vp, C.c_int, _dp], C.c_int),
    "frehg_gw_set_bc_values": ([_vp, C.c_int, _dp], C.c_int),
    "frehg_gw_enable_transport_window": ([_vp], C.c_int),
    "frehg_gw_reset_transport_window": ([_vp], C.c_int),
    "frehg_gw_set_density": ([_vp, _dp, _dp, _dp, _dp, C.c_int], C.c_int),
    "frehg_gw_set_et": ([_vp, _ip, _dp], C.c_int),
    "frehg_gw_attach_coupling": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_int],
                                 C.c_int),
    "frehg_gw_set_timing": ([_vp, C.c_int], C.c_int),
    "frehg_gw_get_timing": ([_vp, C.POINTER(GwTiming)], C.c_int),
    "frehg_gw_synchronize": ([_vp], C.c_int),
    "frehg_gw_set_pcg_shapes": ([_vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "frehg_gw_get_pcg_shapes": ([_vp, _ip, _ip, _ip], C.c_int),
    "frehg_gw_set_step_mode": ([_vp, C.c_int, C.c_int, C.c_int], C.c_int),
    "frehg_gw_get_column_tiles": ([_vp, _ip, _ip, _ip, C.c_int, _ip], C.c_int),
    "frehg_gw_set_column_tiles": ([_vp, _ip, C.c_int], C.c_int),
    "frehg_gw_enable_column_solve": ([_vp, C.c_int], C.c_int),
    "frehg_gw_set_chebyshev": ([_vp, C.c_int, C.c_double], C.c_int),
    "frehg_gw_get_graph_stats": ([_vp, C.POINTER(GraphStats)], C.c_int),
    "frehg_device_malloc": ([C.c_int64, C.POINTER(_vp)], C.c_int),
    "frehg_device_free": ([_vp], C.c_int),
    "frehg_memcpy_htod": ([_vp, _vp, C.c_int64], C.c_int),
    "frehg_memcpy_dtoh": ([_vp, _vp, C.c_int64], C.c_int),
    "frehg_ls_create": ([C.c_int64, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(_vp)],
                        C.c_int),
    "frehg_ls_destroy": ([_vp], C.c_int),
    "frehg_ls_set_pattern": ([_vp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
                             C.c_int),
    "frehg_ls_set_values": ([_vp, _vp], C.c_int),
    "frehg_ls_solve": ([_vp, _vp, _vp, C.POINTER(SolveStats)], C.c_int),
    "frehg_ls_is_symmetric": ([_vp, C.c_double, _ip], C.c_int),
    "frehg_ls_set_precision": ([_vp, C.c_int, C.c_double], C.c_int),
    "frehg_ls_structured": ([_vp, _ip, _dp], C.c_int),
    "frehg_ls_solve_stencil_noalloc": ([C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                        C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.POINTER(SolveStats)], C.c_int),
}

NOALLOC_MAX_CELLS = 1048576  # FREHG_NOALLOC_MAX_CELLS


# The coupling window (include/frehg_coupler.h; src/coupling/Exchange.cpp).
COUPLER_SYMBOLS = {
    "frehg_cpl_begin_window": ([_vp], C.c_int),
    "frehg_cpl_apply_seepage": ([_vp, _vp, C.c_double, _vp, _dp, _dp], C.c_int),
    "frehg_cpl_reduce_audit": ([_vp, _vp, _dp, _dp, _dp], C.c_int),
    "frehg_cpl_refresh_depth": ([_vp, _vp], C.c_int),
}


class HbmProbeResult(C.Structure):
    _fields_ = [("reads", C.c_int), ("writes", C.c_int), ("bytes_per_lane", C.c_int),
                ("nontemporal", C.c_int), ("blocks_per_cu", C.c_int), ("mean_ms", C.c_double),
                ("gbs", C.c_double)]


# The achievable-bandwidth probe (include/frehg_probe.h; no reference counterpart).
PROBE_SYMBOLS = {
    "frehg_hbm_probe": ([C.c_int64, C.c_int, _dp, _dp], C.c_int),
    "frehg_hbm_probe_sweep": ([C.c_int64, C.c_int, C.POINTER(HbmProbeResult), C.c_int, _ip],
                              C.c_int),
    "frehg_probe_last_error": ([], C.c_char_p),
}


# Test exports (include/frehg_gw_testing.h; no reference counterpart): the
# device closures and the post-allocation sweep, for the test suite.
VG_THETA, VG_HEAD, VG_CAPACITY, VG_CONDUCTIVITY, VG_DKDTHETA, VG_CONDUCTIVITY_C = range(6)
TESTING_SYMBOLS = {
    "frehg_test_vg": ([C.c_int, C.c_int64, _dp, _dp, _dp, _dp], C.c_int),
    "frehg_gw_test_post_allocate": ([_vp, C.c_double], C.c_int),
}


class FrehgError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[frehg error {code}] {message}")
        self.code = code


_LIB = None


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load the in-tree HIP library; raise if it is not built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FrehgError(FREHG_EDEVICE,
                         f"HIP library {p} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(p))
    for name, (args, res) in {**SYMBOLS, **COUPLER_SYMBOLS, **PROBE_SYMBOLS,
                             **TESTING_SYMBOLS}.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.frehg_abi_version() != ABI_VERSION:
        raise FrehgError(FREHG_EINVAL, "ABI version mismatch")
    if path is None:
        _LIB = lib
    return lib


def check(lib, rc: int):
    if rc != FREHG_OK:
        msg = lib.frehg_last_error()
        raise FrehgError(rc, msg.decode() if msg else "")


class DeviceArray:
    """A float64 device buffer in the library's HIP runtime (copy in / out)."""

    def __init__(self, host, lib=None):
        import numpy as np
        self.lib = lib or load_library()
        a = np.ascontiguousarray(host, dtype=np.float64)
        self.shape = a.shape
        self.nbytes = a.nbytes
        p = C.c_void_p()
        check(self.lib, self.lib.frehg_device_malloc(self.nbytes, C.byref(p)))
        self.ptr = p
        check(self.lib, self.lib.frehg_memcpy_htod(self.ptr, a.ctypes.data_as(C.c_void_p),
                                                   self.nbytes))

    @classmethod
    def zeros(cls, shape, lib=None):
        import numpy as np
        return cls(np.zeros(shape), lib)

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, dtype=np.float64)
        check(self.lib, self.lib.frehg_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self.ptr,
                                                   self.nbytes))
        return out

    def __del__(self):
        try:
            if self.ptr:
                self.lib.frehg_device_free(self.ptr)
                self.ptr = None
        except Exception:
            pass
