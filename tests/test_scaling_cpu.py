"""The product's host-side scaling (Ruiz then Pock-Chambolle, composed:
skg_ruiz and skg_pock_chambolle_into, src/lp/Scaling.cpp:11-152) against the
oracle's restatement (orc_ruiz, orc_pock_chambolle) on the same matrices,
including empty rows and columns and a 1e-3..1e3 value range."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import sankhya_gpu as sg
import sk_oracle as so


def _csr(rows, cols, M):
    M = sp.csr_matrix(M)
    return so.Csr(rows, cols, M.indptr, M.indices, M.data)


def _matrix(seed, m, n):
    rng = np.random.default_rng(seed)
    M = sp.random(m, n, density=0.02, random_state=seed, format="lil")
    M[m // 3, :] = 0            # an empty row
    M[:, n // 2] = 0            # an empty column
    M = sp.csr_matrix(M)
    M.data = rng.choice([-1.0, 1.0], M.nnz) * 10.0 ** rng.uniform(-3, 3, M.nnz)
    return M


@pytest.mark.parametrize("seed,m,n", [(1, 50, 80), (2, 300, 200), (3, 700, 1500)])
def test_product_scaling_matches_oracle(seed, m, n):
    lib = sg.load_gpu()
    orc = so.load()
    M = _matrix(seed, m, n)
    A, AT = _csr(m, n, M), _csr(n, m, M.T)
    R, Cs = np.zeros(m), np.zeros(n)
    lib.skg_ruiz.restype = None
    lib.skg_ruiz.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, sg.c_double_p, sg.c_double_p]
    lib.skg_pock_chambolle_into.restype = None
    lib.skg_pock_chambolle_into.argtypes = [C.c_void_p, C.c_void_p, sg.c_double_p, sg.c_double_p]
    lib.skg_ruiz(C.addressof(A.s), C.addressof(AT.s), 100, 1e-2, sg.dptr(R), sg.dptr(Cs))
    lib.skg_pock_chambolle_into(C.addressof(A.s), C.addressof(AT.s), sg.dptr(R), sg.dptr(Cs))
    # oracle: Ruiz, scale, Pock-Chambolle on the scaled matrix, compose
    R1, C1 = np.zeros(m), np.zeros(n)
    orc.orc_ruiz(A.ref(), 100, 1e-2, sg.dptr(R1), sg.dptr(C1))
    S = sp.diags(R1) @ M @ sp.diags(C1)
    Ss = _csr(m, n, S)
    R2, C2 = np.zeros(m), np.zeros(n)
    orc.orc_pock_chambolle(Ss.ref(), 1.0, sg.dptr(R2), sg.dptr(C2))
    assert np.allclose(R, R1 * R2, rtol=1e-13, atol=0) and np.allclose(Cs, C1 * C2, rtol=1e-13, atol=0)
    # empty rows and columns keep a unit factor
    assert np.all(R[np.diff(M.indptr) == 0] == 1.0)
    assert np.all(Cs[np.diff(sp.csc_matrix(M).indptr) == 0] == 1.0)
