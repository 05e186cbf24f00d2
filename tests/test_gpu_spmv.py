"""GPU sparse products through the C ABI (sk_gpu_csr_*, sk_gpu_spmv_*).

Known answers follow cuda/tests/spmv_smoke.cu:10-55; the random family follows
tests/cuda/test_gpu_spmv.cpp (GPU vs CPU CSR multiply, 1e-9 there; the
products here are FP64 and held to 1e-12 relative)."""
import ctypes as C

import numpy as np
import pytest

import sankhya_gpu as sg

pytestmark = pytest.mark.gpu

OFF = np.array([0, 2, 4], np.int32)
COL = np.array([0, 2, 1, 2], np.int32)
VAL = np.array([1.0, 2.0, 3.0, 4.0])


def _dev(arr):
    return sg.DeviceArray(arr)


def _sync():
    sg.device_synchronize()
    assert sg.load_gpu().sk_gpu_synchronize() == 0


def test_known_answer_host_convenience():
    lib = sg.load_gpu()
    x = np.array([2.0, 5.0, 7.0])
    y = np.zeros(2)
    assert lib.sk_gpu_spmv_f64(2, 3, 4, sg.iptr(OFF), sg.iptr(COL), sg.dptr(VAL), sg.dptr(x), sg.dptr(y)) == 0
    assert np.array_equal(y, [16.0, 43.0])


def test_rejections_before_allocation():
    lib = sg.load_gpu()
    h = sg.GpuCsr()
    bad_off = np.array([0, 3, 2], np.int32)
    assert lib.sk_gpu_csr_create(C.byref(h), 2, 3, 4, sg.iptr(bad_off), sg.iptr(COL), sg.dptr(VAL)) != 0
    bad_col = np.array([0, 3, 1, 2], np.int32)
    assert lib.sk_gpu_csr_create(C.byref(h), 2, 3, 4, sg.iptr(OFF), sg.iptr(bad_col), sg.dptr(VAL)) != 0
    bad_val = np.array([1.0, np.nan, 3.0, 4.0])
    assert lib.sk_gpu_csr_create(C.byref(h), 2, 3, 4, sg.iptr(OFF), sg.iptr(COL), sg.dptr(bad_val)) != 0
    assert h.device_row_offsets is None


def test_device_chain_known_answers():
    lib = sg.load_gpu()
    h = sg.GpuCsr()
    assert lib.sk_gpu_csr_create(C.byref(h), 2, 3, 4, sg.iptr(OFF), sg.iptr(COL), sg.dptr(VAL)) == 0
    dx = _dev([2.0, 5.0, 7.0])
    dy = _dev(np.zeros(2))
    _sync()
    assert lib.sk_gpu_spmv_device_f64(C.byref(h), dx.value, dy.value) == 0
    _sync()
    assert np.array_equal(dy.numpy(), [16.0, 43.0])
    # A^T [16, 43] = [16, 129, 204]; A [16, 129, 204] = [424, 1203]
    assert lib.sk_gpu_spmv_transpose_device_f64(C.byref(h), dy.value, dx.value) == 0
    assert lib.sk_gpu_spmv_device_f64(C.byref(h), dx.value, dy.value) == 0
    _sync()
    assert np.array_equal(dx.numpy(), [16.0, 129.0, 204.0])
    assert np.array_equal(dy.numpy(), [424.0, 1203.0])
    lib.sk_gpu_csr_destroy(C.byref(h))
    lib.sk_gpu_csr_destroy(C.byref(h))       # idempotent
    assert h.rows == 0 and h.device_row_offsets is None


def _random_csr(rng, m, n, lengths):
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum(lengths)
    cols = np.concatenate([np.sort(rng.choice(n, size=k, replace=False)) for k in lengths]).astype(np.int32) \
        if rp[-1] else np.zeros(0, np.int32)
    vals = rng.standard_normal(rp[-1])
    return rp, cols, vals


@pytest.mark.parametrize("seed,m,n,kind", [
    (7, 100, 80, "uniform"), (11, 1000, 3000, "uniform"), (21, 513, 400, "short"),
    (31, 2000, 5000, "powerlaw"), (41, 50, 4000, "long"), (51, 300, 300, "empty_rows"),
    (52, 1, 10000, "long"),
])
def test_random_family_matches_cpu(seed, m, n, kind):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        lengths = rng.integers(0, min(n, 40), m)
    elif kind == "short":
        lengths = rng.integers(0, 3, m)
    elif kind == "powerlaw":
        lengths = np.minimum(n, (rng.pareto(1.2, m) + 1).astype(int))
    elif kind == "long":
        lengths = rng.integers(500, min(n, 3000), m)
    else:
        lengths = np.where(rng.random(m) < 0.5, 0, rng.integers(1, 20, m))
    rp, cj, v = _random_csr(rng, m, n, lengths)
    x = rng.standard_normal(n)
    import scipy.sparse as sp
    A = sp.csr_matrix((v, cj, rp), shape=(m, n))
    ref = A @ x
    reft = A.T @ (ref / (1.0 + np.abs(ref).max()))
    lib = sg.load_gpu()
    h = sg.GpuCsr()
    assert lib.sk_gpu_csr_create(C.byref(h), m, n, int(rp[-1]), sg.iptr(rp), sg.iptr(cj if len(cj) else np.zeros(1, np.int32)),
                                 sg.dptr(v if len(v) else np.zeros(1))) == 0, sg.last_error()
    dx = _dev(x)
    dy = _dev(np.full(m, np.nan))
    _sync()
    assert lib.sk_gpu_spmv_device_f64(C.byref(h), dx.value, dy.value) == 0
    _sync()
    y = dy.numpy()
    scale = np.abs(A).sum(axis=1).A1 * np.abs(x).max() + 1e-300
    assert np.all(np.abs(y - ref) <= 1e-12 * scale + 1e-300), np.max(np.abs(y - ref) / scale)
    dyin = _dev(ref / (1.0 + np.abs(ref).max()))
    dxo = _dev(np.full(n, np.nan))
    _sync()
    assert lib.sk_gpu_spmv_transpose_device_f64(C.byref(h), dyin.value, dxo.value) == 0
    _sync()
    xt = dxo.numpy()
    scale_t = np.abs(A).sum(axis=0).A1 + 1e-300
    assert np.all(np.abs(xt - reft) <= 1e-12 * scale_t + 1e-300)
    lib.sk_gpu_csr_destroy(C.byref(h))


def test_axpy():
    lib = sg.load_gpu()
    x = np.arange(1000, dtype=np.float64)
    y = np.ones(1000)
    assert lib.sk_gpu_axpy_f64(1000, 2.5, sg.dptr(x), sg.dptr(y)) == 0
    assert np.array_equal(y, np.fma(2.5, np.arange(1000.0), 1.0) if hasattr(np, "fma") else 2.5 * np.arange(1000.0) + 1.0)
    assert lib.sk_gpu_axpy_f64(-1, 1.0, sg.dptr(x), sg.dptr(y)) != 0


def test_gather_rate_probe():
    """The device's random-gather rate (the second ceiling of the sparse
    products) is measured and lands in a physically sensible range."""
    lib = sg.load_gpu()
    rate = C.c_double()
    assert lib.sk_gpu_gather_rate(1 << 20, 1 << 22, 3, C.byref(rate)) == 0, sg.last_error()
    assert 1e9 < rate.value < 1e13
    assert lib.sk_gpu_gather_rate(0, 10, 1, C.byref(rate)) == -1
    assert lib.sk_gpu_gather_rate(10, (1 << 30) + 1, 1, C.byref(rate)) == -1


def test_headline_size_products_match_scipy():
    """Both products at the headline LP's full size (1M x 2M, 15.8M nonzeros,
    Zipf(1.2) column degrees up to 2,000, so A^T has rows long enough for the
    whole-workgroup path) against scipy in FP64, entry by entry within
    1e-12 of each row's absolute mass."""
    w = sg.load_workloads()
    md = sg.Model()
    q = C.c_double()
    assert w.skw_zipf_lp(C.byref(md.m), 1_000_000, 2_000_000, 1.2, 8.0, 2000, 0.3, 20260826, C.byref(q)) == 0
    md._generated = True
    p, i, v = md.csc()
    import scipy.sparse as sp
    A = sp.csc_matrix((v, i, p), shape=(md.nrow, md.ncol)).tocsr()
    A.sort_indices()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(md.ncol)
    y = rng.standard_normal(md.nrow)
    lib = sg.load_gpu()
    h = sg.GpuCsr()
    rp, cj, vals = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    assert lib.sk_gpu_csr_create(C.byref(h), md.nrow, md.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj),
                                 sg.dptr(vals)) == 0, sg.last_error()
    dx, dy = _dev(x), _dev(np.full(md.nrow, np.nan))
    dyin, dxo = _dev(y), _dev(np.full(md.ncol, np.nan))
    _sync()
    assert lib.sk_gpu_spmv_device_f64(C.byref(h), dx.value, dy.value) == 0
    assert lib.sk_gpu_spmv_transpose_device_f64(C.byref(h), dyin.value, dxo.value) == 0
    _sync()
    absA = abs(A)
    scale = absA @ np.abs(x) + 1e-300
    assert np.all(np.abs(dy.numpy() - A @ x) <= 1e-12 * scale)
    scale_t = absA.T @ np.abs(y) + 1e-300
    assert np.all(np.abs(dxo.numpy() - A.T @ y) <= 1e-12 * scale_t)
    lib.sk_gpu_csr_destroy(C.byref(h))
