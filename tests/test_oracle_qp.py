"""Pins of the oracle's QP PDLP (orc_first_order_qp / orc_pdlp_solve_qp).

The reference's GpuPdlp is LP-only; the device engine extends it with a
Hessian and the oracle restates that extension.  It is pinned here by a
zero Hessian (which must reproduce the LP path bit for bit), closed forms,
and the reference C core's own first-order QP solve (golden fixture).
"""
import numpy as np

import sankhya_gpu as sg
import sk_oracle as so
from helpers import golden

INF = sg.SK_INFINITY


def _params(**kw):
    kw.setdefault("precision", sg.PREC_FP64)
    return sg.default_pdlp_params(**kw)


def _dense_model(A, c, rl, ru, l, u, Qd):
    A = np.asarray(A, float)
    Qd = np.asarray(Qd, float)
    m, n = A.shape

    def csc(M):
        p, i, x = [0], [], []
        for j in range(M.shape[1]):
            for r in range(M.shape[0]):
                if M[r, j] != 0:
                    i.append(r); x.append(M[r, j])
            p.append(len(i))
        return p, i, x

    p, i, x = csc(A)
    return sg.Model.from_arrays(m, n, p, i, x, c, rl, ru, l, u, q=csc(Qd))


def test_zero_hessian_reproduces_the_lp_path_bit_for_bit():
    lp = sg.Model.generated("skw_bounded_lp", 200, 400, 5, 11)
    p, i, x = lp.csc()
    vec = [lp.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")]
    qp = sg.Model.from_arrays(lp.nrow, lp.ncol, p, i, x, *vec, q=(np.zeros(lp.ncol + 1), [], []))
    P = _params(max_iterations=2000, eps_optimal=1e-9)
    r1, x1, y1, s1 = so.first_order_pdlp(lp, P)
    r2, x2, y2, s2 = so.first_order_pdlp(qp, P)
    assert r1 == 0 and r2 == 0
    assert (s1.iterations, s1.restarts, s1.rejected_steps) == (s2.iterations, s2.restarts, s2.rejected_steps)
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    assert s1.relative_gap == s2.relative_gap


def test_separable_box_qp_closed_form():
    """min sum d_j x_j^2 / 2 + c_j x_j over a box, one free row: x = clamp(-c/d)."""
    d = np.array([0.5, 1.0, 2.0, 3.0, 4.0, 5.0])
    c = np.array([-1.0, 2.0, -3.0, 0.5, -10.0, 1.0])
    l = np.array([0.0, -1.0, 0.0, -10.0, 0.0, 0.0])
    u = np.array([10.0, 10.0, 1.0, 10.0, 1.0, 10.0])
    model = _dense_model(np.ones((1, 6)), c, [-INF], [INF], l, u, np.diag(d))
    rc, x, y, st = so.first_order_pdlp(model, _params(eps_optimal=1e-10, max_iterations=200000))
    assert rc == 0 and st.converged
    assert np.allclose(x, np.clip(-c / d, l, u), atol=1e-7), x
    assert abs(y[0]) <= 1e-7


def test_equality_constrained_qp_lagrange_system():
    """min (x1^2 + 2 x2^2)/2 - x1 - x2  s.t.  x1 + x2 = 1: the Lagrange system
    gives x = (2/3, 1/3) and, with the reduced cost c + Qx + A'y = 0,
    y = 1/3; objective x'Qx/2 + c'x = -2/3."""
    Qd = np.diag([1.0, 2.0])
    c = np.array([-1.0, -1.0])
    model = _dense_model([[1.0, 1.0]], c, [1.0], [1.0], [-10.0, -10.0], [10.0, 10.0], Qd)
    rc, x, y, st = so.first_order_pdlp(model, _params(eps_optimal=1e-10, max_iterations=200000))
    assert rc == 0 and st.converged
    xs = np.array([2.0 / 3.0, 1.0 / 3.0])
    assert np.allclose(x, xs, atol=1e-7) and abs(y[0] - 1.0 / 3.0) <= 1e-7, (x, y)
    assert abs(0.5 * xs @ Qd @ xs + c @ xs + 2.0 / 3.0) <= 1e-15
    assert abs(st.primal_objective + 2.0 / 3.0) <= 1e-7


def test_matches_the_reference_first_order_qp_solve():
    """The 300 x 600 sparse QP (Q = D + B'B) of the golden fixture, whose
    expected values the reference C core's first-order QP solve wrote:
    the oracle's QP PDLP to 1e-9 lands on the same objective within 1e-8
    relative and the same point (strictly convex: unique) within 1e-7
    (measured: 2.8e-10 and 1.3e-9)."""
    g = golden("qp_small.npz")
    model = sg.Model.generated("skw_sparse_qp", 300, 600, 3, 200, 3, 777)
    rc, x, y, st = so.first_order_pdlp(model, _params(eps_optimal=1e-9, max_iterations=400000))
    assert rc == 0 and st.converged, st.as_dict()
    kkt = so.relative_kkt(model, x, y)
    assert max(kkt[:3]) <= 1.05e-9
    ref = float(g["pdhg_objective"])
    assert abs(kkt[3] - ref) <= 1e-8 * (1.0 + abs(ref)), (kkt[3], ref)
    xr = g["pdhg_x"]
    assert np.linalg.norm(x - xr) <= 1e-7 * (1.0 + np.linalg.norm(xr))


def test_rejects_the_combinations_it_does_not_define():
    model = sg.Model.generated("skw_sparse_qp", 30, 60, 3, 20, 3, 5)
    for kw in (dict(adaptive_step=0), dict(restart_rule=sg.RESTART_NDG)):
        rc, *_ = so.first_order_pdlp(model, _params(max_iterations=64, **kw))
        assert rc == -1, kw
