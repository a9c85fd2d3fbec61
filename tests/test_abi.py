"""The C-ABI library loads without a GPU, exports every function that
include/frehg_gw.h declares, and its ctypes mirror has the C struct layout."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from frehg2_amd import abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "frehg_gw.h"


PROBE_HEADER = ROOT / "include" / "frehg_probe.h"
COUPLER_HEADER = ROOT / "include" / "frehg_coupler.h"
TESTING_HEADER = ROOT / "include" / "frehg_gw_testing.h"


def declared_functions(header=HEADER):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(frehg_\w+)\s*\(", text, re.M)))


def test_header_declares_the_seams():
    names = declared_functions()
    for required in ("frehg_gw_create", "frehg_gw_step", "frehg_gw_get_audit",
                     "frehg_ls_set_pattern", "frehg_ls_solve"):
        assert required in names
    assert set(names) == set(abi.SYMBOLS), set(names) ^ set(abi.SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    for name in (declared_functions() + declared_functions(PROBE_HEADER) +
                 declared_functions(COUPLER_HEADER) + declared_functions(TESTING_HEADER)):
        assert hasattr(lib, name), name
    assert lib.frehg_abi_version() == abi.ABI_VERSION


def test_probe_header_matches_its_ctypes_table():
    assert set(declared_functions(PROBE_HEADER)) == set(abi.PROBE_SYMBOLS)


def test_coupler_header_matches_its_ctypes_table():
    assert set(declared_functions(COUPLER_HEADER)) == set(abi.COUPLER_SYMBOLS)


def test_testing_header_matches_its_ctypes_table():
    assert set(declared_functions(TESTING_HEADER)) == set(abi.TESTING_SYMBOLS)


def test_every_header_is_checked():
    checked = {HEADER, PROBE_HEADER, COUPLER_HEADER, TESTING_HEADER}
    assert set((ROOT / "include").glob("*.h")) <= checked


def test_testing_calls_reject_bad_arguments():
    """Argument checks run before any device work."""
    lib = abi.load_library()
    assert lib.frehg_test_vg(99, 1, None, None, None, None) == abi.FREHG_EINVAL
    assert lib.frehg_gw_test_post_allocate(None, 1.0) != 0


def test_coupler_header_compiles_as_c(tmp_path):
    prog = tmp_path / "cpl.c"
    prog.write_text(f'#include "{COUPLER_HEADER}"\nint main(void){{return 0;}}\n')
    subprocess.run(["gcc", "-Wall", "-Werror", "-o", str(tmp_path / "cpl"), str(prog)], check=True)


def test_coupler_calls_without_coupling_are_rejected():
    """A null solver handle or a solver without coupling is an error code
    (no device work is launched)."""
    lib = abi.load_library()
    assert lib.frehg_cpl_begin_window(None) != 0


def test_probe_rejects_bad_arguments_without_a_device():
    lib = abi.load_library()
    a, b = C.c_double(), C.c_double()
    assert lib.frehg_hbm_probe(0, 5, C.byref(a), C.byref(b)) != 0
    assert b"positive" in lib.frehg_probe_last_error()


def test_ctypes_structs_match_the_c_layout(tmp_path):
    prog = tmp_path / "layout.c"
    fields = {
        "frehg_gw_desc": [n for n, _ in abi.GwDesc._fields_],
        "frehg_gw_audit": [n for n, _ in abi.GwAudit._fields_],
        "frehg_gw_solve_stats": [n for n, _ in abi.SolveStats._fields_],
        "frehg_gw_timing": [n for n, _ in abi.GwTiming._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fl in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fl:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(prog)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True,
                                                        text=True, check=True).stdout.split("\n") if l)
    for st, cls in (("frehg_gw_desc", abi.GwDesc), ("frehg_gw_audit", abi.GwAudit),
                    ("frehg_gw_solve_stats", abi.SolveStats), ("frehg_gw_timing", abi.GwTiming)):
        assert int(out[st]) == C.sizeof(cls), st
        for f, _ in cls._fields_:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_missing_device_is_an_error_not_a_fallback():
    from frehg2_amd.gw import RichardsSolver
    from helpers import gw_base_case
    try:
        RichardsSolver(gw_base_case(1, 1, 4, 0.1, 0.5)).close()
    except abi.FrehgError as e:
        assert e.code == abi.FREHG_EDEVICE
    else:
        pytest.skip("a GPU is visible")


@pytest.mark.gpu
def test_hbm_probe_measures_both_streams():
    from frehg2_amd import hbm
    print("[abi] hbm probe n=4M", flush=True)
    r = hbm.measure(4 * 1024 * 1024, 5)
    # 8 x 32 MB fits the on-package cache, so only sanity is checked here
    assert 100.0 < r["copy_gbs"] < 1e5 and 100.0 < r["stream_6r2w_gbs"] < 1e5, r


@pytest.mark.gpu
def test_step_inputs_reject_what_they_cannot_use():
    # the sink is readable only once set; head and theta need data; the
    # predictor-corrector pair is 0 or 1; a detached sink and a reset
    # density are accepted at any time
    from frehg2_amd import cases
    from frehg2_amd.gw import RichardsSolver
    g = RichardsSolver(cases.loam_32x32x50(steps=1, nx=4, ny=3))
    with pytest.raises(abi.FrehgError) as e:
        g.get(abi.F_SINK)
    assert e.value.code == abi.FREHG_EINVAL
    with pytest.raises(abi.FrehgError) as e:
        abi.check(g.lib, g.lib.frehg_gw_set_field(g._h, abi.F_HEAD, None))
    assert e.value.code == abi.FREHG_EINVAL
    g.set_sink(None)
    g.set_density(None)
    g.set_et(None, None)
    # a channel mask without its flux (or the reverse) is an error, not a detach
    n2 = (g.ny + 2) * (g.nx + 2)
    cols = np.ones(n2, dtype=np.int32)
    q = np.full(n2, 1e-7)
    assert g.lib.frehg_gw_set_et(g._h, cols.ctypes.data_as(abi._ip), None) == abi.FREHG_EINVAL
    assert g.lib.frehg_gw_set_et(g._h, None, q.ctypes.data_as(abi._dp)) == abi.FREHG_EINVAL
    with pytest.raises(abi.FrehgError) as e:
        g.set_et(cols.reshape(g.ny + 2, g.nx + 2), None)
    assert e.value.code == abi.FREHG_EINVAL
    g.step(60.0)
    assert g.audit()["tr_applied"] == 0.0
    g.close()
    c = cases.loam_32x32x50(steps=1, nx=4, ny=3)
    desc, keep = c.to_desc()
    desc.predictor_corrector = 2
    h = abi.C.c_void_p()
    assert g.lib.frehg_gw_create(abi.C.byref(desc), abi.C.byref(h)) == abi.FREHG_EINVAL
