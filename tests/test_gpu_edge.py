"""Degenerate structure on the GPU PDLP engine, against HiGHS on the same
arrays: empty rows and columns, free and one-sided variables, rows with
infinite bounds on one or both sides, duplicate rows, a problem with no
constraints, and the 1e30 sentinel for absent bounds."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

import sankhya_gpu as sg
from helpers import gpu_pdlp

pytestmark = pytest.mark.gpu
PRECS = [sg.PREC_FP64, sg.PREC_MIXED]


def _arrays(A, c, rl, ru, l, u):
    A = sp.csr_matrix(np.asarray(A, dtype=np.float64))
    m, n = A.shape
    vec = {"c": np.asarray(c, float), "rlow": np.asarray(rl, float), "rupp": np.asarray(ru, float),
           "clow": np.asarray(l, float), "cupp": np.asarray(u, float)}
    return m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64), vec


def _highs(A, c, rl, ru, l, u):
    A = np.asarray(A, float)
    def big(v):
        out = v.copy()
        out[v >= 1e30] = np.inf
        out[v <= -1e30] = -np.inf
        return out
    rl, ru, l, u = big(np.asarray(rl, float)), big(np.asarray(ru, float)), big(np.asarray(l, float)), big(np.asarray(u, float))
    fu, fl = np.isfinite(ru), np.isfinite(rl)
    A_ub = np.vstack([A[fu], -A[fl]]) if A.shape[0] else None
    b_ub = np.concatenate([ru[fu], -rl[fl]]) if A.shape[0] else None
    r = linprog(c, A_ub=A_ub if A_ub is not None and A_ub.size else None,
                b_ub=b_ub if b_ub is not None and b_ub.size else None,
                bounds=[(None if not np.isfinite(a) else a, None if not np.isfinite(b) else b) for a, b in zip(l, u)],
                method="highs")
    assert r.status == 0
    return r.fun


CASES = {
    # empty row 2 (0 <= 0 <= 1), empty column 3 (bounded, costs 1), free x1 held by rows
    "empty_row_and_column": ([[1, 1, 0, 0], [1, -1, 0, 0], [0, 0, 0, 0]], [-1, -2, 0, 1],
                             [-np.inf, -1, 0], [4, 2, 1], [0, -np.inf, 0, 0], [10, np.inf, 5, 3]),
    # rows bounded on one side only and a duplicated row, sentinel bounds
    "one_sided_and_duplicate": ([[1, 2, 0], [1, 2, 0], [0, 1, 1], [1, 0, 1]], [-1, -1, -1],
                                [-1e30, -1e30, 1, -1e30], [6, 8, 1e30, 5], [0, 0, 0], [1e30, 1e30, 1e30]),
    # a free row (both sides infinite) next to an equality
    "free_row_and_equality": ([[1, 1, 1], [2, -1, 0]], [1, 2, 3], [3, -np.inf], [3, np.inf],
                              [0, 0, 0], [10, 10, 10]),
}


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("name", sorted(CASES))
def test_degenerate_structure_matches_highs(name, prec):
    A, c, rl, ru, l, u = CASES[name]
    print(f"case {name} prec {prec}", flush=True)
    rc, x, y, st = gpu_pdlp(arrays=_arrays(A, c, rl, ru, l, u), precision=prec, eps_optimal=1e-8)
    assert rc == 0 and st.converged, (rc, sg.last_error())
    ref = _highs(A, c, rl, ru, l, u)
    assert abs(np.dot(c, x) - ref) <= 1e-6 * (1 + abs(ref)), (np.dot(c, x), ref)


@pytest.mark.parametrize("prec", PRECS)
def test_no_constraints(prec):
    """m = 0: the optimum sits at the bound each cost sign picks."""
    c = np.array([1.0, -2.0, 0.5, -1.0])
    l, u = np.array([-1.0, 0.0, 2.0, -3.0]), np.array([4.0, 5.0, 7.0, -1.0])
    arr = (0, 4, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), {
        "c": c, "rlow": np.zeros(0), "rupp": np.zeros(0), "clow": l, "cupp": u})
    rc, x, y, st = gpu_pdlp(arrays=arr, precision=prec, eps_optimal=1e-9)
    assert rc == 0 and st.converged, (rc, sg.last_error())
    assert np.allclose(x, np.where(c > 0, l, u), atol=1e-6)
