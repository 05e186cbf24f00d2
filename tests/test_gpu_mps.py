"""Models read from MPS / QPS files (sk_gpu_read_mps) and solved on the GPU
through sk_gpu_solve, against known optima: unit_lp (x = 1, the reference's
bench/smoke file), tiny_lp (x = (3, 1), -11 by hand), tiny_qp (x = (2, 0),
-4, c/tests/qps_smoke.c), fixed_names (x = (1, 0), 1 by hand)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so

pytestmark = pytest.mark.gpu
MPS = Path(__file__).resolve().parent / "golden" / "mps"


def _solve(model, prec):
    lib = sg.load_gpu()
    o = sg.SkOptions()
    lib.sk_gpu_options_default(C.byref(o))
    g = sg.GpuOptions()
    lib.sk_gpu_solver_options_default(C.byref(g))
    g.pdlp.precision = prec
    s = sg.SkSolution()
    rc = lib.sk_gpu_solve(model.ptr(), C.byref(o), C.byref(g), C.byref(s), None)
    x = np.ctypeslib.as_array(s.x, shape=(model.ncol,)).copy() if s.x else None
    out = (rc, s.result, s.objective, x)
    lib.sk_gpu_solution_free(C.byref(s))
    return out


@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_MIXED])
@pytest.mark.parametrize("name,x_opt,obj", [("unit_lp.mps", [1.0], 1.0),
                                            ("tiny_lp.mps", [3.0, 1.0], -11.0),
                                            ("tiny_qp.qps", [2.0, 0.0], -4.0)])
def test_mps_models_solve_to_known_optimum(name, x_opt, obj, prec):
    model = sg.Model.read_mps(MPS / name)
    print(f"case {name} prec {prec}", flush=True)
    rc, result, objective, x = _solve(model, prec)
    assert rc == 0 and result == sg.SK_RESULT_OPTIMAL, (rc, result, sg.last_error())
    assert abs(objective - obj) <= 1e-5 * (1 + abs(obj))
    assert np.allclose(x, x_opt, atol=1e-4)


def test_fixed_format_model_solves():
    """Names with embedded blanks ("ROW A", "COL 1") need the fixed-format
    pass; min x1 + 2 x2, x1 + x2 >= 1, 2 x1 <= 6, x2 <= 3 -> x = (1, 0), 1."""
    model = sg.Model.read_mps(MPS / "fixed_names.mps")
    a = so.model_arrays(model.m)
    assert a["rowname"].tolist() == ["ROW A", "ROW B"] and a["colname"].tolist() == ["COL 1", "COL 2"]
    rc, result, objective, x = _solve(model, sg.PREC_MIXED)
    assert rc == 0 and result == sg.SK_RESULT_OPTIMAL, (rc, result, sg.last_error())
    assert abs(objective - 1.0) <= 1e-5 and np.allclose(x, [1.0, 0.0], atol=1e-4)
