"""Normalized-duality-gap restarts (restart_rule = SK_GPU_RESTART_NDG).

The device evaluates the NDG of the current and the averaged point in FP64
by a three-stage 32-point search for the trust-region step; the oracle
(orc_ndg, pinned by closed forms, brute force and a saddle point in
tests/test_oracle.py) bisects to rounding.  After the same windows the two
merits agree within the search's resolution (step ratio 2^(1/961), held to
2e-3 relative), and the iterates agree as closely as the KKT-restart engine
does with the oracle.  Termination stays on the relative KKT, so the NDG
engine's results meet the same certificates."""
import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so
from helpers import golden, gpu_pdlp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("windows", [1, 2])
def test_ndg_merits_and_iterates_match_the_oracle(windows):
    """Over the first two windows (one restart) the engine follows the
    oracle's trajectory to 1e-9; later, as with KKT restarts, an
    accept/reject decision of the adaptive step taken on a rounding-level
    difference can send the two on different (equally valid) paths, so the
    merits are compared where the points are the same."""
    model = sg.Model.generated("skw_bounded_lp", 300, 600, 5, 41)
    P = dict(max_iterations=64 * windows, restart_rule=sg.RESTART_NDG, use_graph=0)
    rc_o, xo, yo, sto = so.first_order_pdlp(model, sg.default_pdlp_params(precision=sg.PREC_FP64, **P))
    rc, x, y, st = gpu_pdlp(model, precision=sg.PREC_FP64, **P)
    assert rc == 0 and rc_o == 0 and st.iterations == sto.iterations
    assert st.restarts == sto.restarts, (st.restarts, sto.restarts)
    for a, b in zip(st.restart_merit, sto.restart_merit):
        assert abs(a - b) <= 2e-3 * abs(b) + 1e-12, (list(st.restart_merit), list(sto.restart_merit))
    assert np.linalg.norm(x - xo) <= 1e-6 * (1 + np.linalg.norm(xo))
    assert np.linalg.norm(y - yo) <= 1e-6 * (1 + np.linalg.norm(yo))


@pytest.mark.parametrize("seed", [101, 103, 105])
@pytest.mark.parametrize("prec", [sg.PREC_FP64, sg.PREC_MIXED])
def test_ndg_engine_matches_reference_simplex(seed, prec):
    g = golden(f"small_lp_{seed}.npz")
    model = sg.Model.generated("skw_bounded_lp", int(g["m"]), int(g["n"]), 5, seed)
    rc, x, y, st = gpu_pdlp(model, precision=prec, eps_optimal=1e-7, restart_rule=sg.RESTART_NDG)
    assert rc == 0 and st.converged, sg.last_error()
    ref = float(g["simplex_objective"])
    obj = float(model.vec("c") @ x)
    assert abs(obj - ref) <= 1e-6 * (1 + abs(ref)), (obj, ref)
    kkt = so.relative_kkt(model, x, y)
    assert max(kkt[:3]) <= 1.05e-7, kkt


def test_ndg_rejected_for_a_qp():
    model = sg.Model.generated("skw_sparse_qp", 30, 60, 3, 20, 3, 5)
    import ctypes as C
    lib = sg.load_gpu()
    p, i, v = model.csc()
    rp, cj, vv = sg.csc_to_csr(model.nrow, model.ncol, p, i, v)
    vec = [model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")]
    q = model.q_csc()
    P = sg.default_pdlp_params(restart_rule=sg.RESTART_NDG)
    h = C.c_void_p()
    rc = lib.sk_gpu_pdlp_session_create_qp(model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj),
                                           sg.dptr(vv), sg.iptr(q[0]), sg.iptr(q[1]), sg.dptr(q[2]),
                                           int(q[0][-1]), *[sg.dptr(a) for a in vec], C.byref(P), C.byref(h))
    assert rc == -1 and "NDG" in sg.last_error()
