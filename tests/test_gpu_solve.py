"""sk_gpu_solve / sk_gpu_verify: the sk_solve-style model API on the GPU."""
import ctypes as C

import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so
from helpers import golden, GOLDEN

pytestmark = pytest.mark.gpu


def _solve(model, **params):
    lib = sg.load_gpu()
    o = sg.SkOptions()
    lib.sk_gpu_options_default(C.byref(o))
    g = sg.GpuOptions()
    lib.sk_gpu_solver_options_default(C.byref(g))
    for k, v in params.items():
        setattr(g.pdlp, k, v)
    s = sg.SkSolution()
    st = sg.PdlpStats()
    rc = lib.sk_gpu_solve(model.ptr(), C.byref(o), C.byref(g), C.byref(s), C.byref(st))
    return rc, s, st


def test_gpu_verify_matches_oracle_on_simplex_certificate():
    lib = sg.load_gpu()
    g = golden("small_lp_105.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, 105)
    x, y = g["simplex_x"], g["simplex_y"]
    s = sg.SkSolution()
    s.x, s.y, s.ncol, s.nrow = sg.dptr(x), sg.dptr(y), model.ncol, model.nrow
    assert lib.sk_gpu_verify(model.ptr(), C.byref(s)) == 0
    rc, obj, pinf, dinf, comp = so.verify_orc(model, x, y)
    assert abs(s.objective - obj) <= 1e-12 * (1 + abs(obj))
    assert abs(s.primal_infeasibility - pinf) <= 1e-12
    assert abs(s.dual_infeasibility - dinf) <= 1e-12
    assert abs(s.complementarity - comp) <= 1e-12
    assert abs(s.objective - float(g["simplex_objective"])) <= 1e-12 * (1 + abs(obj))


@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_MIXED])
def test_config0_optimal_against_reference_simplex(prec):
    model = sg.Model.generated("skw_bounded_lp", 2000, 5000, 5, 20260825)
    rc, s, g = _solve(model, precision=prec, eps_optimal=1e-6)
    assert rc == 0 and s.result == sg.SK_RESULT_OPTIMAL, (rc, s.result, sg.last_error())
    if (GOLDEN / "cfg0_simplex.npz").exists():
        ref = float(golden("cfg0_simplex.npz")["simplex_objective"])
    else:
        ref = -5221.5086438670805   # reference simplex, recorded with the fixture script
    assert abs(s.objective - ref) <= 1e-6 * (1 + abs(ref)), (s.objective, ref)
    x = np.ctypeslib.as_array(s.x, shape=(model.ncol,)).copy()
    y = np.ctypeslib.as_array(s.y, shape=(model.nrow,)).copy()
    kkt = so.relative_kkt(model, x, y)
    assert max(kkt[:3]) <= 1e-6
    sg.load_gpu().sk_gpu_solution_free(C.byref(s))


def test_solve_rejects_bad_engine_and_structure():
    lib = sg.load_gpu()
    model = sg.Model.generated("skw_bounded_lp", 20, 40, 5, 1)
    o = sg.SkOptions()
    lib.sk_gpu_options_default(C.byref(o))
    o.lp_engine = 99
    s = sg.SkSolution()
    assert lib.sk_gpu_solve(model.ptr(), C.byref(o), None, C.byref(s), None) == sg.SK_ERR_ARG
