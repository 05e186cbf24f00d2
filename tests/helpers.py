"""Shared helpers for the test-suite: calls through the C ABI and the oracle."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import sankhya_gpu as sg
import sk_oracle as so

GOLDEN = Path(__file__).resolve().parent / "golden"


def model_arrays(model):
    p, i, x = model.csc()
    rp, cj, v = sg.csc_to_csr(model.nrow, model.ncol, p, i, x)
    vec = {k: model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")}
    return rp, cj, v, vec


def gpu_pdlp(model=None, *, arrays=None, x0=None, y0=None, **params):
    """sk_gpu_pdlp_solve through the C ABI; returns rc, x, y, stats."""
    lib = sg.load_gpu()
    if arrays is None:
        m, n = model.nrow, model.ncol
        rp, cj, v, vec = model_arrays(model)
    else:
        m, n, rp, cj, v, vec = arrays
    P = sg.default_pdlp_params(**params)
    x = np.zeros(max(n, 1))
    y = np.zeros(max(m, 1))
    st = sg.PdlpStats()
    x0a = np.ascontiguousarray(x0, np.float64) if x0 is not None else None
    y0a = np.ascontiguousarray(y0, np.float64) if y0 is not None else None
    rc = lib.sk_gpu_pdlp_solve(m, n, int(rp[-1]), sg.iptr(rp), sg.iptr(cj if len(cj) else np.zeros(1, np.int32)),
                               sg.dptr(v if len(v) else np.zeros(1)), None, None, None, 0,
                               sg.dptr(vec["c"]), sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]),
                               sg.dptr(vec["clow"]), sg.dptr(vec["cupp"]),
                               sg.dptr(x0a) if x0a is not None else None,
                               sg.dptr(y0a) if y0a is not None else None,
                               C.byref(P), sg.dptr(x), sg.dptr(y), C.byref(st))
    return rc, x[:n], y[:m], st


def dense_lp(A, c, rl, ru, l, u):
    """Arrays tuple for gpu_pdlp from a dense matrix (tiny hand-made LPs)."""
    A = np.asarray(A, np.float64)
    m, n = A.shape
    rp = [0]
    cj, v = [], []
    for i in range(m):
        for j in range(n):
            if A[i, j] != 0.0:
                cj.append(j)
                v.append(A[i, j])
        rp.append(len(cj))
    vec = dict(c=np.asarray(c, np.float64), rlow=np.asarray(rl, np.float64),
               rupp=np.asarray(ru, np.float64), clow=np.asarray(l, np.float64),
               cupp=np.asarray(u, np.float64))
    return (m, n, np.asarray(rp, np.int32), np.asarray(cj, np.int32), np.asarray(v, np.float64), vec)


def golden(name):
    return np.load(GOLDEN / name, allow_pickle=False)
