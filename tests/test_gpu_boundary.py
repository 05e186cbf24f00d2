"""The drop-in contract of the boundary.

* Device-pointer products order with the null stream as the reference's
  default-stream ABI promises (cuda/include/sankhya_cuda.h:35-37): an
  asynchronous null-stream copy into x before the call is seen, and a
  null-stream hipMemcpy of y after it waits for the product -- with no
  sk_gpu_synchronize in between.
* sk_gpu_solve's certificate gate (c/src/sk_solve.c:987-999): a QP whose
  relative KKT has converged but whose sk_verify residuals miss the
  requested margins is not OPTIMAL; the time limit covers every refinement
  round.
* Cooperative cancel (src/cuda/GpuPdlp.hpp:60-66, GpuPdlp.cpp:430-434).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import sankhya_gpu as sg
import sk_oracle as so
from helpers import gpu_pdlp

pytestmark = pytest.mark.gpu


def test_device_products_order_with_the_null_stream():
    lib = sg.load_gpu()
    hip = sg._hip_runtime()
    m, n, per_row = 400_000, 800_000, 16
    g = np.random.default_rng(4)
    cols = g.integers(0, n, size=(m, per_row))
    A = sp.csr_matrix((g.standard_normal(m * per_row), cols.ravel(), np.arange(0, m * per_row + 1, per_row)),
                      shape=(m, n))
    A.sum_duplicates()
    A.sort_indices()
    rp, cj = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    h = sg.GpuCsr()
    assert lib.sk_gpu_csr_create(C.byref(h), m, n, int(rp[-1]), sg.iptr(rp), sg.iptr(cj), sg.dptr(A.data)) == 0
    dx = sg.DeviceArray(np.full(n, np.nan))
    dy = sg.DeviceArray(np.full(m, np.nan))
    dxt = sg.DeviceArray(np.full(n, np.nan))
    sg.device_synchronize()
    rng = np.random.default_rng(9)
    y = np.empty(m)
    xt = np.empty(n)
    for rep in range(4):
        x = rng.standard_normal(n)
        # null stream: async H2D into x, then the library's product, then a
        # blocking D2H of y -- no device or library synchronisation between
        assert hip.hipMemcpyAsync(dx.value, x.ctypes.data, 8 * n, 1, None) == 0
        assert lib.sk_gpu_spmv_device_f64(C.byref(h), dx.value, dy.value) == 0
        assert lib.sk_gpu_spmv_transpose_device_f64(C.byref(h), dy.value, dxt.value) == 0
        assert hip.hipMemcpy(y.ctypes.data, dy.value, 8 * m, 2) == 0
        assert hip.hipMemcpy(xt.ctypes.data, dxt.value, 8 * n, 2) == 0
        ref = A @ x
        assert np.all(np.abs(y - ref) <= 1e-12 * (abs(A) @ np.abs(x) + 1e-300)), rep
        ref_t = A.T @ ref
        assert np.all(np.abs(xt - ref_t) <= 1e-11 * (abs(A).T @ np.abs(ref) + 1e-300)), rep
    lib.sk_gpu_csr_destroy(C.byref(h))


def _qp():
    return sg.Model.generated("skw_sparse_qp", 300, 600, 3, 200, 3, 777)


def _solve(model, o_over=None, **pdlp):
    lib = sg.load_gpu()
    o = sg.SkOptions()
    lib.sk_gpu_options_default(C.byref(o))
    for k, v in (o_over or {}).items():
        setattr(o, k, v)
    g = sg.GpuOptions()
    lib.sk_gpu_solver_options_default(C.byref(g))
    for k, v in pdlp.items():
        setattr(g.pdlp, k, v)
    s = sg.SkSolution()
    st = sg.PdlpStats()
    assert lib.sk_gpu_solve(model.ptr(), C.byref(o), C.byref(g), C.byref(s), C.byref(st)) == 0
    x = np.ctypeslib.as_array(s.x, shape=(model.ncol,)).copy()
    y = np.ctypeslib.as_array(s.y, shape=(model.nrow,)).copy()
    out = dict(result=s.result, pinf=s.primal_infeasibility, dinf=s.dual_infeasibility,
               comp=s.complementarity, seconds=s.solve_seconds, x=x, y=y, stats=st)
    lib.sk_gpu_solution_free(C.byref(s))
    return out


def test_qp_converged_kkt_but_failed_certificate_is_not_optimal():
    model = _qp()
    r = _solve(model, dict(primal_tol=1e-15, dual_tol=1e-15), eps_optimal=1e-7)
    # the returned point meets the relative KKT tolerance (the oracle's
    # evaluation) ...
    kkt = so.relative_kkt(model, r["x"], r["y"])
    assert max(kkt[:3]) <= 1e-6, kkt
    # ... but sk_verify's residuals exceed 100 x 1e-15, so the result is
    # the reference's unconverged first-order outcome, not OPTIMAL
    assert max(r["dinf"], r["comp"], r["pinf"]) > 1e-13
    assert r["result"] == sg.SK_RESULT_ITERATION_LIMIT, r["result"]
    _, obj, pinf, dinf, comp = so.verify_orc(model, r["x"], r["y"])
    assert abs(pinf - r["pinf"]) <= 1e-12 and abs(dinf - r["dinf"]) <= 1e-12


def test_time_limit_covers_every_round():
    model = sg.Model.generated("skw_bounded_lp", 2000, 5000, 5, 20260825)
    r = _solve(model, dict(time_limit=0.05), eps_optimal=1e-9)
    assert r["result"] == sg.SK_RESULT_TIME_LIMIT, r["result"]
    assert r["seconds"] < 0.05 + 2.0


@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_MIXED])
def test_cancel_flag_stops_at_the_next_window(prec):
    model = sg.Model.generated("skw_bounded_lp", 200, 400, 5, 3)
    flag = C.c_int(1)
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-12, cancel=C.pointer(flag))
    assert rc == 0 and st.cancelled == 1 and st.converged == 0
    assert st.iterations == 64 and np.all(np.isfinite(x)) and np.all(np.isfinite(y))
    flag.value = 0
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-7, cancel=C.pointer(flag))
    assert rc == 0 and st.cancelled == 0 and st.converged == 1


@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_MIXED])
def test_session_resume_continues_the_same_solve(prec):
    """A solve to 1e-4 resumed with a 1e-7 tolerance reaches 1e-7 (the
    oracle's FP64 relative KKT of the point) in about the iterations of a
    cold solve to 1e-7: it continues the same iterate, step size and primal
    weight instead of starting over."""
    lib = sg.load_gpu()
    model = sg.Model.generated("skw_bounded_lp", 400, 800, 5, 11)
    m, n = model.nrow, model.ncol
    p, i, v = model.csc()
    rp, cj, vv = sg.csc_to_csr(m, n, p, i, v)
    vec = [model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")]

    def session(**kw):
        P = sg.default_pdlp_params(precision=prec, **kw)
        h = C.c_void_p()
        assert lib.sk_gpu_pdlp_session_create(m, n, int(rp[-1]), sg.iptr(rp), sg.iptr(cj), sg.dptr(vv),
                                              *[sg.dptr(a) for a in vec], C.byref(P), C.byref(h)) == 0
        return h

    x, y = np.zeros(n), np.zeros(m)
    st = sg.PdlpStats()
    h = session(eps_optimal=1e-4)
    assert lib.sk_gpu_pdlp_session_solve(h, None, None, sg.dptr(x), sg.dptr(y), C.byref(st)) == 0
    assert st.converged
    first = st.iterations
    assert lib.sk_gpu_pdlp_session_set_limits(h, 500000, 0.0, 1e-7) == 0
    st2 = sg.PdlpStats()
    assert lib.sk_gpu_pdlp_session_resume(h, sg.dptr(x), sg.dptr(y), C.byref(st2)) == 0
    assert st2.converged and st2.iterations > first
    kkt = so.relative_kkt(model, x, y)
    assert max(kkt[:3]) <= 1.05e-7, kkt
    stc = sg.PdlpStats()
    assert lib.sk_gpu_pdlp_session_solve(h, None, None, sg.dptr(x), sg.dptr(y), C.byref(stc)) == 0
    assert stc.converged
    assert st2.iterations <= 1.5 * stc.iterations + 640, (st2.iterations, stc.iterations)
    lib.sk_gpu_pdlp_session_destroy(h)
    # nothing to resume on a fresh session
    h = session()
    assert lib.sk_gpu_pdlp_session_resume(h, sg.dptr(x), sg.dptr(y), C.byref(st2)) == -1
    lib.sk_gpu_pdlp_session_destroy(h)
