"""Plain PDHG solver family (sk_gpu_lp_pdhg, ..._qp_pdhg_*): the known answers
and rejection cases of cuda/tests/pdhg_smoke.cu and diagonal_qp_smoke.cu,
and agreement with the CPU restatement of cuda/src/sankhya_cuda_lp.cu."""
import ctypes as C

import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so
from helpers import model_arrays

pytestmark = pytest.mark.gpu
INF = np.inf


def _csr(lib, rows, cols, rp, cj, v):
    h = sg.GpuCsr()
    assert lib.sk_gpu_csr_create(C.byref(h), rows, cols, int(rp[-1]), sg.iptr(np.asarray(rp, np.int32)),
                                 sg.iptr(np.asarray(cj, np.int32)), sg.dptr(np.asarray(v, np.float64))) == 0
    return h


def _a(*v):
    return np.asarray(v, np.float64)


def test_pdhg_smoke_known_answer_and_rejections():
    lib = sg.load_gpu()
    h = _csr(lib, 1, 1, [0, 1], [0], [1.0])
    st = sg.LpSettings(5000, 10, 0.25, 2.0, 1.0, 1e-5, 0.0)
    res = sg.LpResult()
    sol = _a(0.0)
    # NaN cost rejected
    assert lib.sk_gpu_lp_pdhg(C.byref(h), sg.dptr(_a(np.nan)), sg.dptr(_a(1.0)), sg.dptr(_a(INF)),
                              sg.dptr(_a(0.0)), sg.dptr(_a(INF)), st, sg.dptr(sol), C.byref(res)) != 0
    # overflow in the product is a numeric failure (-2), never convergence
    ho = _csr(lib, 1, 1, [0, 1], [0], [1.0e308])
    rc = lib.sk_gpu_lp_pdhg(C.byref(ho), sg.dptr(_a(1.0)), sg.dptr(_a(1.0e308)), sg.dptr(_a(INF)),
                            sg.dptr(_a(0.0)), sg.dptr(_a(INF)), st, sg.dptr(sol), C.byref(res))
    assert rc == -2 and res.status == -2
    lib.sk_gpu_csr_destroy(C.byref(ho))
    rc = lib.sk_gpu_lp_pdhg(C.byref(h), sg.dptr(_a(1.0)), sg.dptr(_a(1.0)), sg.dptr(_a(INF)),
                            sg.dptr(_a(0.0)), sg.dptr(_a(INF)), st, sg.dptr(sol), C.byref(res))
    assert rc == 0 and res.status == 0
    assert res.maximum_row_violation <= 1e-5 and res.maximum_kkt_residual <= 1e-5
    assert res.maximum_dual_residual <= 1e-5 and res.maximum_complementarity <= 1e-5
    assert abs(sol[0] - 1.0) <= 1e-3
    lib.sk_gpu_csr_destroy(C.byref(h))


def test_diagonal_qp_smoke():
    lib = sg.load_gpu()
    h = _csr(lib, 1, 1, [0, 1], [0], [1.0])
    args = [sg.dptr(_a(2.0)), sg.dptr(_a(-4.0)), sg.dptr(_a(1.0)), sg.dptr(_a(INF)), sg.dptr(_a(0.0)), sg.dptr(_a(INF))]
    st = sg.LpSettings(10000, 10, 0.15, 0.5, 1.0, 1e-7, 0.0)
    res = sg.LpResult()
    sol = _a(0.0)
    rc = lib.sk_gpu_diagonal_qp_pdhg(C.byref(h), *args, st, sg.dptr(sol), C.byref(res))
    assert rc == 0 and res.status == 0 and abs(sol[0] - 2.0) <= 1e-5 and abs(res.objective + 4.0) <= 1e-5
    assert res.maximum_kkt_residual <= 1e-7 and res.maximum_dual_residual <= 1e-7
    vs = sg.LpSettings(10000, 10, 0.0, 0.0, 1.0, 1e-7, 0.0)
    dual = _a(0.0)
    sol[0] = 0.0
    rc = lib.sk_gpu_qp_pdhg_preconditioned_with_dual(C.byref(h), None, *args, sg.dptr(_a(0.15)), sg.dptr(_a(0.5)),
                                                     vs, sg.dptr(sol), sg.dptr(dual), C.byref(res))
    assert rc == 0 and res.status == 0 and abs(sol[0] - 2.0) <= 1e-5 and abs(dual[0]) <= 1e-6
    rc = lib.sk_gpu_qp_pdhg_preconditioned(C.byref(h), None, *args, sg.dptr(_a(0.0)), sg.dptr(_a(0.5)),
                                           vs, sg.dptr(sol), C.byref(res))
    assert rc != 0
    lib.sk_gpu_csr_destroy(C.byref(h))


def test_matches_cpu_restatement_on_seeded_lp():
    lib = sg.load_gpu()
    model = sg.Model.generated("skw_bounded_lp", 40, 80, 5, 3)
    rp, cj, v, vec = model_arrays(model)
    h = _csr(lib, model.nrow, model.ncol, rp, cj, v)
    st = sg.LpSettings(20000, 20, 0.0, 0.0, 1.0, 1e-6, 0.0)
    absA = np.abs(v)
    col_mass = np.bincount(cj, weights=absA, minlength=model.ncol)
    row_mass = np.add.reduceat(absA, rp[:-1]) * (np.diff(rp) > 0)
    ps = 0.9 / (1.0 + col_mass)
    ds = 0.9 / (1.0 + row_mass)
    order = ("c", "rlow", "rupp", "clow", "cupp")
    x = np.zeros(model.ncol); y = np.zeros(model.nrow); res = sg.LpResult()
    rc = lib.sk_gpu_qp_pdhg_preconditioned_with_dual(C.byref(h), None, None, *[sg.dptr(vec[k]) for k in order],
                                                     sg.dptr(ps), sg.dptr(ds), st, sg.dptr(x), sg.dptr(y), C.byref(res))
    assert rc == 0
    A = so.Csr(model.nrow, model.ncol, rp, cj, v)
    xo = np.zeros(model.ncol); yo = np.zeros(model.nrow); ro = sg.LpResult()
    orc = so.load()
    rc_o = orc.orc_cuda_pdhg(A.ref(), None, None, *[sg.dptr(vec[k]) for k in order], st, sg.dptr(ps),
                             sg.dptr(ds), sg.dptr(xo), sg.dptr(yo), C.byref(ro))
    assert rc_o == 0
    assert res.status == ro.status
    assert abs(res.iterations - ro.iterations) <= st.check_every
    c = vec["c"]
    assert abs(c @ x - c @ xo) <= 1e-6 * (1 + abs(c @ xo))
    lib.sk_gpu_csr_destroy(C.byref(h))
