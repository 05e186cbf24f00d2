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


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("use_graph", [0, 1])
def test_fused_decision_matches_separate_launch(prec, use_graph):
    """The step decision taken inside the next dual prox (every workgroup
    reducing the probe partials itself) gives the same iterates, bit for bit,
    as the one-workgroup decision launch."""
    model = sg.Model.generated("skw_bounded_lp", 400, 900, 5, 17)
    a = gpu_pdlp(model, eps_optimal=1e-7, precision=prec, use_graph=use_graph, unfused=1)
    b = gpu_pdlp(model, eps_optimal=1e-7, precision=prec, use_graph=use_graph, unfused=0)
    assert a[0] == 0 and b[0] == 0 and a[3].converged and b[3].converged
    assert a[3].iterations == b[3].iterations and a[3].restarts == b[3].restarts
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert b[3].kernel_launches < a[3].kernel_launches


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


def _qp_arrays(A, c, rl, ru, l, u, Q):
    arr = dense_lp(A, c, rl, ru, l, u)
    Qd = np.asarray(Q, float)
    rp, cj, v = [0], [], []
    for i in range(Qd.shape[0]):
        for j in range(Qd.shape[1]):
            if Qd[i, j] != 0:
                cj.append(j); v.append(Qd[i, j])
        rp.append(len(cj))
    return arr, (np.asarray(rp, np.int32), np.asarray(cj, np.int32), np.asarray(v, float))


def _gpu_qp(arr, q, **params):
    import ctypes as C
    lib = sg.load_gpu()
    m, n, rp, cj, v, vec = arr
    P = sg.default_pdlp_params(**params)
    x = np.zeros(n); y = np.zeros(max(m, 1)); st = sg.PdlpStats()
    rc = lib.sk_gpu_pdlp_solve(m, n, int(rp[-1]), sg.iptr(rp), sg.iptr(cj), sg.dptr(v),
                               sg.iptr(q[0]), sg.iptr(q[1]), sg.dptr(q[2]), int(q[0][-1]),
                               sg.dptr(vec["c"]), sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]),
                               sg.dptr(vec["clow"]), sg.dptr(vec["cupp"]), None, None, C.byref(P),
                               sg.dptr(x), sg.dptr(y), C.byref(st))
    return rc, x, y[:m], st


@pytest.mark.parametrize("prec", PRECS)
def test_qp_closed_forms(prec):
    # general_qp_active_smoke.c: min x'Qx/2 - 3x1 - 3x2, Q=[[2,1],[1,2]], x1+x2<=1 -> (1/2,1/2), -2.25
    arr, q = _qp_arrays([[1, 1]], [-3, -3], [-INF], [1], [0, 0], [10, 10], [[2, 1], [1, 2]])
    rc, x, y, st = _gpu_qp(arr, q, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged, sg.last_error()
    assert np.allclose(x, [0.5, 0.5], atol=1e-6)
    # diagonal_qp_smoke.c (cuda): min x^2 - 4x st x >= 1 -> x = 2, objective -4, y = 0
    arr, q = _qp_arrays([[1]], [-4], [1], [INF], [0], [INF], [[2]])
    rc, x, y, st = _gpu_qp(arr, q, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged
    assert abs(x[0] - 2.0) < 1e-6 and abs(y[0]) < 1e-6


def test_qp_reference_first_order_fixture():
    g = golden("qp_small.npz")
    model = sg.Model.generated("skw_sparse_qp", 300, 600, 3, 200, 3, 777)
    import ctypes as C
    lib = sg.load_gpu()
    o = sg.SkOptions(); lib.sk_gpu_options_default(C.byref(o))
    go = sg.GpuOptions(); lib.sk_gpu_solver_options_default(C.byref(go))
    go.pdlp.eps_optimal = 1e-7
    s = sg.SkSolution()
    assert lib.sk_gpu_solve(model.ptr(), C.byref(o), C.byref(go), C.byref(s), None) == 0
    assert s.result == sg.SK_RESULT_OPTIMAL, sg.last_error()
    x = np.ctypeslib.as_array(s.x, shape=(model.ncol,)).copy()
    y = np.ctypeslib.as_array(s.y, shape=(model.nrow,)).copy()
    # the reference's CPU QP PDHG stops near 1e-5 KKT: agree with it to that level
    ref = float(g["pdhg_objective"])
    assert abs(s.objective - ref) <= 1e-4 * (1 + abs(ref)), (s.objective, ref)
    # and certify this point independently (convex QP: KKT residuals)
    rc, obj, pinf, dinf, comp = so.verify_orc(model, x, y)
    kkt = so.relative_kkt(model, x, y)
    assert max(kkt[:3]) <= 1e-6, kkt
    assert pinf <= 1e-5
    lib.sk_gpu_solution_free(C.byref(s))


@pytest.mark.parametrize("prec", PRECS)
def test_qp_fused_decision_matches_separate_launch(prec):
    """A QP's step probe runs Q x and A x in one launch (two tile sets);
    with the decision fused into the next dual prox the iterates are the
    same, bit for bit, as with the one-workgroup decision launch, and the
    solution satisfies the relative KKT tolerance against the oracle's
    evaluation."""
    model = sg.Model.generated("skw_sparse_qp", 300, 600, 3, 200, 3, 777)
    p, i, v = model.csc()
    rp, cj, vv = sg.csc_to_csr(model.nrow, model.ncol, p, i, v)
    vec = {k: model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")}
    arr = (model.nrow, model.ncol, rp, cj, vv, vec)
    q = model.q_csc()
    a = _gpu_qp(arr, q, eps_optimal=1e-7, precision=prec, unfused=1)
    b = _gpu_qp(arr, q, eps_optimal=1e-7, precision=prec, unfused=0)
    assert a[0] == 0 and b[0] == 0 and a[3].converged and b[3].converged, sg.last_error()
    assert a[3].iterations == b[3].iterations and a[3].restarts == b[3].restarts
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert b[3].kernel_launches < a[3].kernel_launches
    kkt = so.relative_kkt(model, b[1], b[2])
    assert max(kkt[:3]) <= 1e-6, kkt


def test_headline_size_first_iterations_match_the_oracle():
    """The FP64 engine on the headline LP at its full size (1M x 2M, 15.8M
    nonzeros) against the oracle's FP64 PDLP (Ruiz + Pock-Chambolle scaling,
    30 power iterations, adaptive steps), both stopped after 8 iterations:
    the same ||A|| and the same iterate, up to the summation order of the
    products and reductions."""
    import ctypes as C
    w = sg.load_workloads()
    model = sg.Model()
    q = C.c_double()
    assert w.skw_zipf_lp(C.byref(model.m), 1_000_000, 2_000_000, 1.2, 8.0, 2000, 0.3, 20260826, C.byref(q)) == 0
    model._generated = True
    P = dict(max_iterations=8, precision=sg.PREC_FP64, use_graph=0)
    rc_o, xo, yo, sto = so.first_order_pdlp(model, sg.default_pdlp_params(**P))
    rc, x, y, st = gpu_pdlp(model, **P)
    assert rc == 0 and rc_o == 0 and st.iterations == 8 and sto.iterations == 8
    assert abs(st.matrix_norm - sto.matrix_norm) <= 1e-10 * sto.matrix_norm
    assert np.linalg.norm(x - xo) <= 1e-8 * (1.0 + np.linalg.norm(xo))
    assert np.linalg.norm(y - yo) <= 1e-8 * (1.0 + np.linalg.norm(yo))


@pytest.mark.parametrize("prec", PRECS)
def test_autotuned_session_solves_and_reports_its_choice(prec):
    """autotune = 1 times each candidate tiling on the device and keeps the
    fastest; whatever it keeps, the solve reaches the reference simplex
    objective, and the stats name the choice."""
    g = golden("small_lp_105.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, 105)
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-7, autotune=1)
    assert rc == 0 and st.converged, sg.last_error()
    assert st.tuned == 1 and st.tile_nnz in (256, 512, 1024, 2048) and st.spmv_grid in (1024, 4096)
    assert st.probe_grid in (512, 1024, 4096) and st.xcache_bytes in (0, 32768, 65536)
    assert st.stream_loads in (0, 1) and 0.0 < st.tune_seconds < 5.0
    ref = float(g["simplex_objective"])
    assert abs(float(model.vec("c") @ x) - ref) <= 1e-6 * (1 + abs(ref))
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-7)
    assert st.tuned == 0 and st.spmv_grid == 4096


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_cols", [5, 998, 1003])
def test_x_copy_in_lds_edges_reach_the_simplex_objective(prec, n_cols):
    """The step probe's LDS copy of x when it covers every column (and
    fewer than one 16-byte chunk of them, or a column count that is not a
    multiple of a chunk, whose last columns are then gathered from memory):
    the solve reaches the reference simplex objective."""
    m = max(3, n_cols // 2)
    model = sg.Model.generated("skw_bounded_lp", m, n_cols, 3, 4242 + n_cols)
    T = dict(autotune=0, tile_nnz=256, spmv_grid=4096, probe_grid=512, xcache_bytes=65536, stream_loads=1)
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-7, **T)
    assert rc == 0 and st.converged and st.xcache_bytes == 65536, sg.last_error()
    rc0, x0, y0, st0 = gpu_pdlp(model, precision=prec, eps_optimal=1e-7)
    assert rc0 == 0 and st0.converged and st0.xcache_bytes == 0
    f, f0 = float(model.vec("c") @ x), float(model.vec("c") @ x0)
    assert abs(f - f0) <= 1e-5 * (1 + abs(f0))


def test_invalid_lds_copy_sizes_are_rejected():
    """A fixed tiling's LDS copy of x must be a multiple of 16 bytes and at
    most 65,536: anything else is an error at session creation, not a
    launch that could read past the copy."""
    model = sg.Model.generated("skw_bounded_lp", 40, 100, 3, 7)
    for xc in (100, 65536 + 16, -16):
        rc, x, y, st = gpu_pdlp(model, autotune=0, tile_nnz=256, xcache_bytes=xc)
        assert rc != 0 and "xcache_bytes" in sg.last_error()
    rc, x, y, st = gpu_pdlp(model, autotune=0, tile_nnz=256, probe_grid=5000)
    assert rc != 0 and "probe_grid" in sg.last_error()


@pytest.mark.parametrize("tile", [256, 512, 1024, 2048])
def test_every_tiling_matches_the_oracle_on_a_power_law_lp(tile):
    """The products' tile paths at every entry cap the autotuner can pick:
    a 20,000 x 40,000 instance of the power-law generator (column degrees up
    to 2,000, so A^T has rows longer than every cap -- the workgroup-wide
    long-row path -- next to tiles of one to hundreds of rows, and 1, 2, 4
    or 8 entries per lane), both grid caps, and a step probe on its own
    grid cap gathering x's longest columns (4,096 to 16,384 of the 40,000,
    by precision and copy size) from an LDS copy.  The FP64 engine after 8 adaptive
    iterations against the oracle's FP64 PDLP within summation-order
    rounding; the mixed engine (FP32 products) within the FP32 drift of 8
    iterations."""
    import ctypes as C
    w = sg.load_workloads()
    model = sg.Model()
    q = C.c_double()
    assert w.skw_zipf_lp(C.byref(model.m), 20_000, 40_000, 1.2, 8.0, 2000, 0.3, 20260826, C.byref(q)) == 0
    model._generated = True
    P = dict(max_iterations=8, use_graph=0)
    rc_o, xo, yo, sto = so.first_order_pdlp(model, sg.default_pdlp_params(precision=sg.PREC_FP64, **P))
    assert rc_o == 0 and sto.iterations == 8
    for grid, probe, xc in ((1024, 1024, 0), (4096, 4096, 0), (4096, 512, 65536), (1024, 1024, 32768)):
        T = dict(autotune=0, tile_nnz=tile, spmv_grid=grid, probe_grid=probe, xcache_bytes=xc,
                 stream_loads=1 + (grid == 4096))
        rc, x, y, st = gpu_pdlp(model, precision=sg.PREC_FP64, **P, **T)
        assert rc == 0 and st.iterations == 8, sg.last_error()
        assert st.tile_nnz == tile and st.spmv_grid == grid and st.probe_grid == probe
        assert st.xcache_bytes == xc
        assert abs(st.matrix_norm - sto.matrix_norm) <= 1e-10 * sto.matrix_norm
        assert np.linalg.norm(x - xo) <= 1e-9 * (1.0 + np.linalg.norm(xo))
        assert np.linalg.norm(y - yo) <= 1e-9 * (1.0 + np.linalg.norm(yo))
        rc, xm, ym, st = gpu_pdlp(model, precision=sg.PREC_MIXED, **P, **T)
        assert rc == 0 and st.iterations == 8 and st.fp32_iterations == 8, sg.last_error()
        assert np.linalg.norm(xm - xo) <= 1e-4 * (1.0 + np.linalg.norm(xo))
        assert np.linalg.norm(ym - yo) <= 1e-4 * (1.0 + np.linalg.norm(yo))
