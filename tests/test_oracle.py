"""The CPU oracle, pinned: bit-for-bit against the reference C core
(oracle/_ref, where it was built) and against the committed golden fixtures
written from it; closed forms and invariants for what those leave open."""
import ctypes as C
import hashlib

import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so
from helpers import golden, dense_lp

SMALL = [101, 102, 103, 104, 105]
INF = np.inf


def digest(model):
    h = hashlib.sha256()
    for a in model.csc():
        h.update(np.ascontiguousarray(a).tobytes())
    for k in ("c", "rlow", "rupp", "clow", "cupp"):
        h.update(model.vec(k).tobytes())
    q = model.q_csc()
    if q is not None:
        for a in q:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("seed", SMALL)
def test_generator_reproduces_fixture_inputs(seed):
    g = golden(f"small_lp_{seed}.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, seed)
    assert digest(model) == str(g["digest"])


@pytest.mark.parametrize("seed", SMALL)
def test_cpu_pdhg_matches_reference_fixture_bitwise(seed):
    g = golden(f"small_lp_{seed}.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, seed)
    rc, s = so.first_order_orc(model, int(g["pdhg_limit"]))
    x, y = so.solution_arrays(s)
    assert rc == int(g["pdhg_rc"]) and s.result == int(g["pdhg_result"])
    assert s.iterations == int(g["pdhg_iterations"])
    assert np.array_equal(x, g["pdhg_x"]) and np.array_equal(y, g["pdhg_y"])
    assert s.objective == float(g["pdhg_objective"])


def test_cfg0_pdhg_trajectory_bitwise():
    g = golden("cfg0_pdhg.npz")
    model = sg.Model.generated("skw_bounded_lp", 2000, 5000, 5, 20260825)
    assert digest(model) == str(g["digest"])
    rc, s = so.first_order_orc(model, 400)
    x, y = so.solution_arrays(s)
    assert np.array_equal(x, g["pdhg_x"]) and np.array_equal(y, g["pdhg_y"])
    assert s.iterations == int(g["pdhg_iterations"]) and s.result == int(g["pdhg_result"])


def test_qp_first_order_path_bitwise():
    g = golden("qp_small.npz")
    model = sg.Model.generated("skw_sparse_qp", 300, 600, 3, 200, 3, 777)
    assert digest(model) == str(g["digest"])
    rc, s = so.first_order_orc(model, 20000)
    x, y = so.solution_arrays(s)
    assert rc == int(g["pdhg_rc"]) and s.result == int(g["pdhg_result"])
    assert s.iterations == int(g["pdhg_iterations"])
    assert np.array_equal(x, g["pdhg_x"]) and np.array_equal(y, g["pdhg_y"])


@pytest.mark.parametrize("seed", SMALL)
def test_verify_matches_reference_on_simplex_certificate(seed):
    g = golden(f"small_lp_{seed}.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, seed)
    rc, obj, pinf, dinf, comp = so.verify_orc(model, g["simplex_x"], g["simplex_y"])
    assert rc == 0
    assert obj == float(g["simplex_objective"])
    assert pinf == float(g["simplex_pinf"]) and dinf == float(g["simplex_dinf"])
    assert comp == float(g["simplex_comp"])


def test_reference_build_bitwise_when_present():
    if so.load_ref() is None:
        pytest.skip("oracle/_ref not built here (reference sources absent)")
    for m, n, seed, lim in [(30, 60, 5, 3000), (150, 300, 6, 500)]:
        model = sg.Model.generated("skw_bounded_lp", m, n, 5, seed)
        r1, s1 = so.first_order_ref(model, lim)
        r2, s2 = so.first_order_orc(model, lim)
        x1, y1 = so.solution_arrays(s1)
        x2, y2 = so.solution_arrays(s2)
        assert r1 == r2 and s1.result == s2.result and s1.iterations == s2.iterations
        assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


def _tiny_model(A, c, rl, ru, l, u):
    A = np.asarray(A, float)
    m, n = A.shape
    p = [0]; i = []; x = []
    for j in range(n):
        for r in range(m):
            if A[r, j] != 0:
                i.append(r); x.append(A[r, j])
        p.append(len(i))
    return sg.Model.from_arrays(m, n, p, i, x, c, rl, ru, l, u)


def test_verify_closed_forms():
    # min x st x >= 1 (row), x >= 0: optimum x = 1 with y = -1 (rc = c + A'y = 0)
    model = _tiny_model([[1.0]], [1.0], [1.0], [sg.SK_INFINITY], [0.0], [sg.SK_INFINITY])
    rc, obj, pinf, dinf, comp = so.verify_orc(model, [1.0], [-1.0])
    assert (obj, pinf, dinf, comp) == (1.0, 0.0, 0.0, 0.0)
    rc, obj, pinf, dinf, comp = so.verify_orc(model, [0.5], [-1.0])
    assert pinf == 0.5
    # wrong multiplier sign at the lower activity bound is a dual infeasibility
    rc, obj, pinf, dinf, comp = so.verify_orc(model, [1.0], [1.0])
    assert dinf == 2.0      # rc = 2 on an interior-free variable, and y > 0 at rlow
    kkt = so.relative_kkt(model, np.array([1.0]), np.array([-1.0]))
    assert kkt[0] == 0.0 and kkt[2] == 0.0


def test_relative_kkt_primal_residual_closed_form():
    model = _tiny_model([[1.0, 1.0]], [1.0, 1.0], [5.0], [5.0], [0.0, 0.0], [10.0, 10.0])
    kkt = so.relative_kkt(model, np.array([1.0, 1.0]), np.array([0.0]))
    # ||Ax - proj(Ax)|| = 3, ||b|| = sqrt(25 + 25)
    assert abs(kkt[0] - 3.0 / (1.0 + np.sqrt(50.0))) < 1e-15


def test_ruiz_and_pock_chambolle_invariants():
    orc = so.load()
    rng = np.random.default_rng(3)
    import scipy.sparse as sp
    A = sp.random(40, 60, density=0.1, random_state=4, format="csr")
    A.data = np.sign(rng.standard_normal(A.nnz)) * 10 ** rng.uniform(-3, 3, A.nnz)
    M = so.Csr(40, 60, A.indptr, A.indices, A.data)
    R = np.zeros(40); Cc = np.zeros(60)
    orc.orc_ruiz(M.ref(), 100, 1e-2, sg.dptr(R), sg.dptr(Cc))
    S = sp.diags(R) @ A @ sp.diags(Cc)
    rn = abs(S).max(axis=1).toarray().ravel()
    cn = abs(S).max(axis=0).toarray().ravel()
    if not (np.all(R == 1) and np.all(Cc == 1)):
        assert np.all(np.abs(rn[rn > 0] - 1) < 1e-2) and np.all(np.abs(cn[cn > 0] - 1) < 1e-2)
    Z = so.Csr(3, 4, np.zeros(4, np.int32), [], [])
    R3 = np.zeros(3); C4 = np.zeros(4)
    orc.orc_ruiz(Z.ref(), 100, 1e-2, sg.dptr(R3), sg.dptr(C4))
    assert np.all(R3 == 1) and np.all(C4 == 1)
    # Pock-Chambolle alpha=1 on [[1, 3], [0, 4]]: R = 1/sqrt(row abs sums), C = 1/sqrt(col abs sums)
    P = so.Csr(2, 2, [0, 2, 3], [0, 1, 1], [1.0, 3.0, 4.0])
    R2 = np.zeros(2); C2 = np.zeros(2)
    orc.orc_pock_chambolle(P.ref(), 1.0, sg.dptr(R2), sg.dptr(C2))
    assert np.allclose(R2, [1 / 2.0, 1 / 2.0], rtol=0, atol=1e-16)
    assert np.allclose(C2, [1.0, 1 / np.sqrt(7.0)], rtol=0, atol=1e-16)


def test_power_norm_closed_forms():
    orc = so.load()
    D = so.Csr(3, 3, [0, 1, 2, 3], [0, 1, 2], [3.0, 1.0, 2.0])
    assert abs(orc.orc_power_norm(D.ref(), 60) - 3.0) < 1e-9
    u = np.array([1.0, 2.0, 2.0]); v = np.array([3.0, 4.0])
    R1 = np.outer(u, v)
    M = so.Csr(3, 2, [0, 2, 4, 6], [0, 1, 0, 1, 0, 1], R1.ravel())
    assert abs(orc.orc_power_norm(M.ref(), 30) - 15.0) < 1e-12
    rng = np.random.default_rng(8)
    A = rng.standard_normal((30, 20))
    Ms = so.Csr(30, 20, np.arange(0, 601, 20), np.tile(np.arange(20), 30), A.ravel())
    est = orc.orc_power_norm(Ms.ref(), 30)
    true = np.linalg.norm(A, 2)
    assert est <= true * (1 + 1e-12) and est >= 0.9 * true


@pytest.mark.parametrize("case", ["textbook", "equality", "range", "bounds"])
def test_oracle_pdlp_hand_optima(case):
    lp = {
        "textbook": ([[1, 0], [0, 2], [3, 2]], [-3, -5], [-INF] * 3, [4, 12, 18], [0, 0], [INF, INF], [2, 6]),
        "equality": ([[1, 1]], [1, 1], [5], [5], [0, 0], [10, 10], None),
        "range": ([[1]], [-1], [2], [7], [0], [100], [7]),
        "bounds": ([[1, 1]], [-1, -1], [-INF], [100], [0, 0], [3, 4], [3, 4]),
    }[case]
    big = lambda v: [sg.SK_INFINITY if b == INF else (-sg.SK_INFINITY if b == -INF else b) for b in v]
    model = _tiny_model(lp[0], lp[1], big(lp[2]), big(lp[3]), big(lp[4]), big(lp[5]))
    rc, x, y, st = so.first_order_pdlp(model, sg.default_pdlp_params(eps_optimal=1e-9))
    assert rc == 0 and st.converged
    if lp[6] is not None:
        assert np.allclose(x, lp[6], atol=1e-6)
    else:
        assert abs(x.sum() - 5) < 1e-6


@pytest.mark.parametrize("seed", SMALL)
def test_oracle_pdlp_matches_reference_simplex(seed):
    g = golden(f"small_lp_{seed}.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, seed)
    rc, x, y, st = so.first_order_pdlp(model, sg.default_pdlp_params(eps_optimal=1e-8))
    assert rc == 0 and st.converged
    ref = float(g["simplex_objective"])
    assert abs(model.vec("c") @ x - ref) <= 1e-6 * (1 + abs(ref))
    assert max(so.relative_kkt(model, x, y)[:3]) <= 1e-6


def test_cuda_pdhg_restatement_known_answers():
    orc = so.load()
    A = so.Csr(1, 1, [0, 1], [0], [1.0])
    a = lambda *v: np.asarray(v, float)
    st = sg.LpSettings(5000, 10, 0.25, 2.0, 1.0, 1e-5, 0.0)
    x = a(0.0); r = sg.LpResult()
    args = [sg.dptr(a(1.0)), sg.dptr(a(1.0)), sg.dptr(a(INF)), sg.dptr(a(0.0)), sg.dptr(a(INF))]
    assert orc.orc_cuda_pdhg(A.ref(), None, None, *args, st, None, None, sg.dptr(x), None, C.byref(r)) == 0
    assert r.status == 0 and abs(x[0] - 1.0) < 1e-3
    bad = [sg.dptr(a(np.nan))] + args[1:]
    assert orc.orc_cuda_pdhg(A.ref(), None, None, *bad, st, None, None, sg.dptr(x), None, C.byref(r)) != 0
    O = so.Csr(1, 1, [0, 1], [0], [1e308])
    ov = [sg.dptr(a(1.0)), sg.dptr(a(1e308)), sg.dptr(a(INF)), sg.dptr(a(0.0)), sg.dptr(a(INF))]
    assert orc.orc_cuda_pdhg(O.ref(), None, None, *ov, st, None, None, sg.dptr(x), None, C.byref(r)) == -2
    st2 = sg.LpSettings(10000, 10, 0.15, 0.5, 1.0, 1e-7, 0.0)
    qd = a(2.0)
    args2 = [sg.dptr(a(-4.0)), sg.dptr(a(1.0)), sg.dptr(a(INF)), sg.dptr(a(0.0)), sg.dptr(a(INF))]
    assert orc.orc_cuda_pdhg(A.ref(), None, sg.dptr(qd), *args2, st2, None, None, sg.dptr(x), None, C.byref(r)) == 0
    assert r.status == 0 and abs(x[0] - 2.0) < 1e-5 and abs(r.objective + 4.0) < 1e-5
    assert orc.orc_cuda_pdhg(A.ref(), None, sg.dptr(qd), *args2, st2, sg.dptr(a(0.0)), sg.dptr(a(0.5)),
                             sg.dptr(x), None, C.byref(r)) != 0


def test_validation_semantics():
    orc = so.load()
    model = _tiny_model([[1.0]], [1.0], [1.0], [2.0], [0.0], [1.0])
    assert orc.orc_model_validate(model.ptr()) == 0
    bad = _tiny_model([[1.0]], [np.nan], [1.0], [2.0], [0.0], [1.0])
    assert orc.orc_model_validate(bad.ptr()) == sg.SK_ERR_STRUCTURE
    crossed = _tiny_model([[1.0]], [1.0], [3.0], [2.0], [0.0], [1.0])
    assert orc.orc_model_validate(crossed.ptr()) == sg.SK_ERR_STRUCTURE


# ---------------------------------------------------------------- NDG pins

def test_ndg_closed_form_one_variable():
    """min 2x, x in [0, 10], omega = 4: from x = 3 with restart point 5 the
    ball of radius sqrt(4)*2 = 4 allows a step of 2 toward the cheaper side,
    G = 2 * 2, merit 4 / 4; from x = 1 (restart 5, r = 8) the bound stops
    the step at 0, G = 2 * 1, merit 2 / 8."""
    import numpy as np
    A = np.zeros((0, 1))
    e = np.zeros(0)
    v = so.ndg(A, [2.0], [0.0], [10.0], e, e, [3.0], e, [5.0], e, 4.0)
    assert abs(v - 1.0) <= 1e-12, v
    v = so.ndg(A, [2.0], [0.0], [10.0], e, e, [1.0], e, [5.0], e, 4.0)
    assert abs(v - 0.25) <= 1e-12, v
    # no movement since the restart point: 0
    assert so.ndg(A, [2.0], [0.0], [10.0], e, e, [5.0], e, [5.0], e, 4.0) == 0.0


def test_ndg_is_the_maximum_over_the_ball_brute_force():
    """A 2 x 2 LP: the merit times r bounds the gap of every sampled point of
    the box/ball intersection and is nearly attained by one of them."""
    import numpy as np
    A = np.array([[1.0, 2.0], [3.0, -1.0]])
    c = np.array([1.0, -2.0])
    l, u = np.array([0.0, -1.0]), np.array([4.0, 3.0])
    rl, ru = np.array([-1.0, -np.inf]), np.array([2.0, 5.0])
    x, y = np.array([1.0, 0.5]), np.array([0.3, -0.2])
    xr, yr = np.array([1.6, 0.1]), np.array([-0.1, 0.2])
    om = 1.7
    v = so.ndg(A, c, l, u, rl, ru, x, y, xr, yr, om)
    r = np.sqrt(om * np.sum((x - xr) ** 2) + np.sum((y - yr) ** 2) / om)

    def supp(yv):
        return np.where(yv > 0, np.where(np.isfinite(ru), yv * ru, 0.0),
                        np.where(yv < 0, np.where(np.isfinite(rl), yv * rl, 0.0), 0.0)).sum(-1)

    def gap(xh, yh):   # L(x, yh) - L(xh, y), L = c'x + y'Ax - p(y)
        return (c @ x + yh @ (A @ x) - supp(yh)) - (xh @ c + (xh @ A.T) @ y - supp(y))

    rng = np.random.default_rng(0)
    d = rng.standard_normal((400000, 4))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= rng.uniform(0, 1, (len(d), 1)) ** 0.25 * r
    xh = x + d[:, :2] / np.sqrt(om)
    yh = y + d[:, 2:] * np.sqrt(om)
    ok = np.all((xh >= l) & (xh <= u), axis=1)
    g = gap(xh[ok], yh[ok])
    assert g.max() <= v * r * (1 + 1e-12) + 1e-12
    assert g.max() >= 0.97 * v * r, (g.max(), v * r)


def test_ndg_vanishes_at_a_saddle_point():
    """The textbook LP of tests/lp/test_pdlp.cpp:53-70 (max 3x + 5y, optimum
    (2, 6)) with its optimal multipliers: no point of any ball has a positive
    gap, so the merit is 0 whatever the restart point."""
    import numpy as np
    A = np.array([[1.0, 0.0], [0.0, 2.0], [3.0, 2.0]])
    c = np.array([-3.0, -5.0])
    l, u = np.zeros(2), np.full(2, np.inf)
    rl, ru = np.full(3, -np.inf), np.array([4.0, 12.0, 18.0])
    xs = np.array([2.0, 6.0])
    ys = np.array([0.0, 1.5, 1.0])   # c + A'y = 0, y >= 0 on the active rows
    assert np.allclose(c + A.T @ ys, 0)
    for xr, yr in ((np.zeros(2), np.zeros(3)), (np.array([3.0, 1.0]), np.array([1.0, 0.0, 2.0]))):
        v = so.ndg(A, c, l, u, rl, ru, xs, ys, xr, yr, 2.0)
        assert abs(v) <= 1e-12, v
