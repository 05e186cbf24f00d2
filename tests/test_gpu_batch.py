"""Batched PDLP (one persistent launch, one workgroup per instance): each
instance agrees with the CPU oracle's PDLP and with an independent LP
solver, and its result does not depend on the batch it was solved in."""
import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so

pytestmark = pytest.mark.gpu


def _models(count, lo, hi, seed):
    w = sg.load_workloads()
    sizes = (sg.C.c_int * count)()
    assert w.skw_batch_sizes(count, lo, hi, seed, sizes) == 0
    return [sg.Model.generated("skw_bounded_lp", sizes[i], 2 * sizes[i], 5, seed * count + i) for i in range(count)]


def test_batch_matches_oracle_and_highs():
    models = _models(48, 20, 120, 11)
    rc, xs, ys, res, st = sg.batch_solve(models, eps_optimal=1e-7, max_iterations=200000)
    assert rc == 0, sg.last_error()
    assert st.converged == len(models)
    from scipy.optimize import linprog
    import scipy.sparse as sp
    for k, md in enumerate(models):
        c = md.vec("c")
        p, i, x = md.csc()
        A = sp.csc_matrix((x, i, p), shape=(md.nrow, md.ncol))
        r = linprog(c, A_ub=sp.vstack([A, -A]), b_ub=np.concatenate([md.vec("rupp"), -md.vec("rlow")]),
                    bounds=list(zip(md.vec("clow"), md.vec("cupp"))), method="highs")
        assert r.status == 0
        assert abs(c @ xs[k] - r.fun) <= 1e-6 * (1 + abs(r.fun)), (k, c @ xs[k], r.fun)
        assert abs(res[k].objective - c @ xs[k]) <= 1e-9 * (1 + abs(r.fun))
        kkt = so.relative_kkt(md, xs[k], ys[k])
        assert max(kkt[:3]) <= 1e-5, (k, kkt)
    # the CPU oracle's PDLP reaches the same optimum
    rc_o, xo, yo, sto = so.first_order_pdlp(models[0], sg.default_pdlp_params(eps_optimal=1e-8))
    assert abs(models[0].vec("c") @ xo - models[0].vec("c") @ xs[0]) <= 1e-6 * (1 + abs(models[0].vec("c") @ xo))


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_instance_result_independent_of_batch(threads):
    models = _models(24, 20, 80, 5)
    rc, xs, ys, res, st = sg.batch_solve(models, block_threads=threads, eps_optimal=1e-6)
    assert rc == 0 and st.block_threads == threads
    for k in (0, 7, 23):
        rc1, x1, y1, r1, _ = sg.batch_solve([models[k]], block_threads=threads, eps_optimal=1e-6)
        assert rc1 == 0
        assert r1[0].iterations == res[k].iterations
        assert np.array_equal(x1[0], xs[k]) and np.array_equal(y1[0], ys[k])


def test_graph_and_direct_launch_agree():
    models = _models(16, 20, 60, 3)
    a = sg.batch_solve(models, block_threads=128, use_graph=0, eps_optimal=1e-6)
    b = sg.batch_solve(models, block_threads=128, use_graph=1, eps_optimal=1e-6)
    assert a[0] == 0 and b[0] == 0
    for k in range(len(models)):
        assert np.array_equal(a[1][k], b[1][k])


def test_iteration_limit_reported():
    models = _models(8, 50, 60, 9)
    rc, xs, ys, res, st = sg.batch_solve(models, eps_optimal=1e-9, max_iterations=128)
    assert rc == 0
    assert all(r.iterations == 128 and r.status == 1 for r in res)


@pytest.mark.parametrize("refine", [0, 1])
def test_fp64_tail_and_fp64_mode_agree(refine):
    """Plain FP32 iterates (refine 0) hand the instances that reach
    tail_switch to the FP64 launch; anchored FP32 iterates (refine 1, the
    default) solve every instance to 1e-8 with no FP64 tail.  FP64 iterates
    throughout reach the same optimum."""
    models = _models(16, 20, 100, 21)
    rc_m, xm, ym, rm, sm = sg.batch_solve(models, eps_optimal=1e-8, max_iterations=200000,
                                          precision=sg.PREC_MIXED, refine=refine)
    rc_d, xd, yd, rd, sd = sg.batch_solve(models, eps_optimal=1e-8, max_iterations=200000,
                                          precision=sg.PREC_FP64)
    assert rc_m == 0 and rc_d == 0, sg.last_error()
    assert sm.converged == len(models) and sd.converged == len(models)
    if refine:
        assert sm.handoffs == 0 and sm.fp32_iterations == sm.total_iterations
    else:
        assert sm.handoffs > 0 and sm.fp32_iterations > 0
    assert sd.fp32_iterations == 0 and sd.handoffs == 0
    for k, md in enumerate(models):
        c = md.vec("c")
        assert abs(c @ xm[k] - c @ xd[k]) <= 1e-6 * (1 + abs(c @ xd[k])), k
        assert 0 < rm[k].fp32_iterations <= rm[k].iterations
        assert max(so.relative_kkt(md, xm[k], ym[k])[:3]) <= 1e-6


def test_full_size_batch_sample_meets_the_tolerance_in_the_oracle():
    """The config-3 batch at its full size (4,096 LPs, m ~ U{20..400}, the
    benchmark's seed, autotuned block size): every instance reports
    convergence (eps 1e-6 on the scaled instance), and a spread sample of 64
    returned points has the oracle's FP64 relative KKT of the unscaled
    instance within 1e-5 (the bar the smaller test above uses), with the
    reported objective c'x."""
    count, seed = 4096, 20260828
    models = _models(count, 20, 400, seed)
    rc, xs, ys, res, st = sg.batch_solve(models, eps_optimal=1e-6, max_iterations=200000)
    assert rc == 0, sg.last_error()
    assert st.converged == count and all(r.status == 0 for r in res)
    assert st.block_threads in (128, 256, 512)
    for k in range(0, count, count // 64):
        md = models[k]
        kkt = so.relative_kkt(md, xs[k], ys[k])
        assert max(kkt[:3]) <= 1e-5, (k, kkt)
        c = md.vec("c")
        assert abs(res[k].objective - c @ xs[k]) <= 1e-9 * (1 + abs(c @ xs[k])), k
