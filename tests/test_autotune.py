"""The PCG launch shapes (frehg_gw_set_pcg_shapes) and the shape tuner
(frehg2_amd.autotune): the shapes change the order of partial sums only, so
a tuned solver still meets the oracle at the parity tolerance."""
import json

import numpy as np
import pytest

from frehg2_amd import abi, autotune, cases


class _FakeSolver:
    def __init__(self):
        self.case = cases.loam_32x32x50(steps=1)
        self.shapes = {"march64": 16, "march32": 64, "column": 16}

    def set_pcg_shapes(self, a, b, c):
        self.shapes = {"march64": a, "march32": b, "column": c}


def test_shape_cache_round_trip(tmp_path):
    g = _FakeSolver()
    p = tmp_path / "shapes.json"
    assert autotune.apply(g, p) is None
    autotune.save(p, g, {"march64": 8, "march32": 32, "column": 16}, device="dev")
    data = json.loads(p.read_text())
    assert list(data) == ["dev:32x32x50:fp64:zline"]
    assert autotune.apply(g, p, device="dev") == {"march64": 8, "march32": 32, "column": 16}
    assert g.shapes == {"march64": 8, "march32": 32, "column": 16}


@pytest.mark.gpu
def test_shapes_set_get_and_reject_bad_values():
    from frehg2_amd.gw import RichardsSolver
    # nx even, 64 < nz <= 256: the direction kernels run over column pairs, 4 waves a CU
    deep = RichardsSolver(cases.hetero_512x512x100(nx=16, ny=10, nz=100))
    assert deep.pcg_shapes() == {"march64": 4, "march32": 4, "column": 16}
    deep.close()
    # nz = 50 (one row a lane fills the waves better) or nx odd: one column a wave
    g = RichardsSolver(cases.loam_32x32x50(steps=1))
    assert g.pcg_shapes() == {"march64": 16, "march32": 32, "column": 16}
    odd = RichardsSolver(cases.hetero_512x512x100(nx=15, ny=10, nz=100))
    assert odd.pcg_shapes() == {"march64": 16, "march32": 32, "column": 16}
    odd.close()
    g.set_pcg_shapes(8, 0, 32)
    assert g.pcg_shapes() == {"march64": 8, "march32": 32, "column": 32}
    with pytest.raises(abi.FrehgError):
        g.set_pcg_shapes(-1, 0, 0)
    with pytest.raises(abi.FrehgError):
        g.set_pcg_shapes(0, 300, 0)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tuned_solver_meets_the_oracle(precision):
    """Tune, then three steps against the oracle's (FP64 1e-10 relative head;
    mixed against the FP64 oracle 1e-8); tuning leaves the state as it was."""
    from frehg2_amd.gw import RichardsSolver
    from oracle.oracle import OracleRichards
    case = cases.loam_32x32x50(steps=3, nx=24, ny=20)
    case.precision = precision
    g = RichardsSolver(case)
    h0 = g.get(abi.F_HEAD)
    res = autotune.tune(g, dt=case.dt, repeats=1,
                        candidates={"march64": (8, 32), "march32": (32, 128), "column": (8, 32)})
    assert np.array_equal(g.get(abi.F_HEAD), h0)
    sh = res["shapes"]
    assert g.pcg_shapes() == sh and len(res["trials_ms"]) >= 4
    case64 = cases.loam_32x32x50(steps=3, nx=24, ny=20)
    o = OracleRichards(case64)
    for _ in range(3):
        g.step(case.dt)
        o.step(case.dt)
    hg, ho = g.get(abi.F_HEAD)[1:-1, 1:-1], o.get(abi.F_HEAD)[1:-1, 1:-1]
    tol = 1e-10 if precision == "fp64" else 1e-8
    assert np.max(np.abs(hg - ho)) <= tol * np.max(np.abs(ho))
    g.close()
