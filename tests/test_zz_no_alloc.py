"""sk_gpu_relative_kkt_device_noalloc against sk_gpu_relative_kkt.

The allocation-free entry point reduces the KKT terms inside the products'
epilogues into a static device scratch array; the allocating one uploads the
model, forms A x and c + A^T y, then reduces.  Same FP64 quantities in a
different summation order: held to 1e-12 relative.  Two calls in a row on
the same GPU, the second on a smaller LP, so stale scratch from the first
call would show."""
import ctypes as C

import numpy as np
import pytest

import sankhya_gpu as sg

pytestmark = pytest.mark.gpu


def _reference(model, x, y):
    lib = sg.load_gpu()
    p, i, v = model.csc()
    rp, cj, vv = sg.csc_to_csr(model.nrow, model.ncol, p, i, v)
    vec = [model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")]
    out = np.zeros(5)
    assert lib.sk_gpu_relative_kkt(model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj), sg.dptr(vv),
                                   *[sg.dptr(a) for a in vec], sg.dptr(x), sg.dptr(y), sg.dptr(out)) == 0
    return out, (rp, cj, vv, vec)


def _noalloc(model, arrays, x, y):
    lib = sg.load_gpu()
    rp, cj, vv, vec = arrays
    h = sg.GpuCsr()
    assert lib.sk_gpu_csr_create(C.byref(h), model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj),
                                 sg.dptr(vv)) == 0, sg.last_error()
    dev = [sg.DeviceArray(a) for a in vec] + [sg.DeviceArray(x), sg.DeviceArray(y)]
    dout = sg.DeviceArray(np.full(5, np.nan))
    sg.device_synchronize()
    rc = lib.sk_gpu_relative_kkt_device_noalloc(C.byref(h), *[d.value for d in dev], dout.value)
    assert rc == 0, sg.last_error()
    out = dout.numpy()
    lib.sk_gpu_csr_destroy(C.byref(h))
    return out


def _point(model, seed):
    rng = np.random.default_rng(seed)
    x = np.clip(rng.uniform(-1, 11, model.ncol), model.vec("clow"), model.vec("cupp"))
    y = rng.standard_normal(model.nrow)
    return x, y


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_kkt_matches_the_allocating_entry_point(order):
    for (m, n, seed) in ((2000, 5000, 20260825), (120, 240, 7)):
        model = sg.Model.generated("skw_bounded_lp", m, n, 5, seed)
        x, y = _point(model, seed)
        ref, arrays = _reference(model, x, y)
        got = _noalloc(model, arrays, x, y)
        assert np.all(np.isfinite(got)), got
        assert np.all(np.abs(got - ref) <= 1e-12 * (np.abs(ref) + 1.0)), (m, n, got, ref)


def test_noalloc_kkt_rejects_null_pointers():
    lib = sg.load_gpu()
    model = sg.Model.generated("skw_bounded_lp", 20, 40, 5, 3)
    p, i, v = model.csc()
    rp, cj, vv = sg.csc_to_csr(model.nrow, model.ncol, p, i, v)
    h = sg.GpuCsr()
    assert lib.sk_gpu_csr_create(C.byref(h), model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj),
                                 sg.dptr(vv)) == 0
    assert lib.sk_gpu_relative_kkt_device_noalloc(C.byref(h), *([None] * 8)) == -1
    assert "null" in sg.last_error()
    lib.sk_gpu_csr_destroy(C.byref(h))
