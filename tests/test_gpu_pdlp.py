"""GPU PDLP engine (sk_gpu_pdlp_solve) against hand-computed optima
(tests/lp/test_pdlp.cpp:53-150), the reference simplex (golden fixtures) and
the CPU oracle.  Tolerance: objective within 1e-6 relative and the PDLP
relative KKT triple <= 1e-6, as the product contract states."""
import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so
from helpers import dense_lp, golden, gpu_pdlp

pytestmark = pytest.mark.gpu
INF = np.inf
PRECS = [sg.PREC_FP64, sg.PREC_MIXED]


@pytest.mark.parametrize("prec", PRECS)
def test_textbook_lp(prec):
    # max 3x + 5y st x <= 4, 2y <= 12, 3x + 2y <= 18, x, y >= 0 -> (2, 6), -36
    arr = dense_lp([[1, 0], [0, 2], [3, 2]], [-3, -5], [-INF] * 3, [4, 12, 18], [0, 0], [INF, INF])
    rc, x, y, st = gpu_pdlp(arrays=arr, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged, sg.last_error()
    assert abs(x[0] - 2) < 1e-5 and abs(x[1] - 6) < 1e-5
    assert st.host_syncs < st.iterations / 8 + 16


@pytest.mark.parametrize("prec", PRECS)
def test_equality_row(prec):
    arr = dense_lp([[1, 1]], [1, 1], [5], [5], [0, 0], [10, 10])
    rc, x, y, st = gpu_pdlp(arrays=arr, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged
    assert abs(x.sum() - 5) < 1e-6


@pytest.mark.parametrize("prec", PRECS)
def test_row_range(prec):
    arr = dense_lp([[1]], [-1], [2], [7], [0], [100])
    rc, x, y, st = gpu_pdlp(arrays=arr, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged
    assert abs(x[0] - 7) < 1e-6


@pytest.mark.parametrize("prec", PRECS)
def test_variable_bounds(prec):
    arr = dense_lp([[1, 1]], [-1, -1], [-INF], [100], [0, 0], [3, 4])
    rc, x, y, st = gpu_pdlp(arrays=arr, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged
    assert abs(x[0] - 3) < 1e-6 and abs(x[1] - 4) < 1e-6


@pytest.mark.parametrize("seed", [101, 102, 103, 104, 105])
@pytest.mark.parametrize("prec", PRECS)
def test_small_seeds_match_reference_simplex(seed, prec):
    g = golden(f"small_lp_{seed}.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, seed)
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-7)
    assert rc == 0 and st.converged, sg.last_error()
    ref = float(g["simplex_objective"])
    obj = float(model.vec("c") @ x)
    assert abs(obj - ref) <= 1e-6 * (1 + abs(ref)), (obj, ref)
    kkt = so.relative_kkt(model, x, y)
    assert max(kkt[:3]) <= 1e-6, kkt
    _, _, pinf, dinf, comp = so.verify_orc(model, x, y)
    assert pinf <= 1e-6 * (1 + np.abs(model.vec("rupp")).max())


def test_gpu_matches_cpu_oracle_pdlp():
    model = sg.Model.generated("skw_bounded_lp", 200, 400, 5, 105)
    P = sg.default_pdlp_params(eps_optimal=1e-8)
    rc_o, xo, yo, sto = so.first_order_pdlp(model, P)
    rc, x, y, st = gpu_pdlp(model, eps_optimal=1e-8)
    assert rc == 0 and rc_o == 0 and st.converged and sto.converged
    c = model.vec("c")
    assert abs(c @ x - c @ xo) <= 1e-6 * (1 + abs(c @ xo))
    assert abs(st.matrix_norm - sto.matrix_norm) <= 1e-12 * sto.matrix_norm


def test_bitwise_reproducible():
    model = sg.Model.generated("skw_bounded_lp", 300, 600, 5, 9)
    r1 = gpu_pdlp(model, eps_optimal=1e-7)
    r2 = gpu_pdlp(model, eps_optimal=1e-7)
    assert r1[3].iterations == r2[3].iterations
    assert np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])


@pytest.mark.parametrize("prec", PRECS)
def test_warm_start_from_solution(prec):
    model = sg.Model.generated("skw_bounded_lp", 200, 400, 5, 104)
    rc, x, y, st = gpu_pdlp(model, eps_optimal=1e-7)
    assert rc == 0 and st.converged
    rc, x2, y2, st2 = gpu_pdlp(model, x0=x, y0=y, eps_optimal=1e-6, precision=prec)
    assert rc == 0 and st2.converged
    assert st2.iterations <= st.iterations // 4


def test_graph_window_matches_direct_launches():
    model = sg.Model.generated("skw_bounded_lp", 200, 400, 5, 103)
    a = gpu_pdlp(model, eps_optimal=1e-7, use_graph=0)
    b = gpu_pdlp(model, eps_optimal=1e-7, use_graph=1)
    assert a[3].iterations == b[3].iterations
    assert np.array_equal(a[1], b[1])


def test_fixed_step_path():
    model = sg.Model.generated("skw_bounded_lp", 50, 100, 5, 102)
    rc, x, y, st = gpu_pdlp(model, eps_optimal=1e-6, adaptive_step=0)
    assert rc == 0 and st.converged
    g = golden("small_lp_102.npz")
    ref = float(g["simplex_objective"])
    assert abs(model.vec("c") @ x - ref) <= 1e-5 * (1 + abs(ref))


def test_iteration_limit_and_rejection():
    model = sg.Model.generated("skw_bounded_lp", 50, 100, 5, 102)
    rc, x, y, st = gpu_pdlp(model, max_iterations=10)
    assert rc == 0 and not st.converged and st.iterations == 10
    arr = dense_lp([[1]], [np.nan], [1], [2], [0], [1])
    rc, *_ = gpu_pdlp(arrays=arr)
    assert rc == -1
