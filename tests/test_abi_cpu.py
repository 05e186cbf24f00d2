"""Host-side checks that need no GPU: the C ABI library loads and exports
every entry point include/sankhya_gpu.h declares; argument validation happens
before any device call; the workload generators follow their laws."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import sankhya_gpu as sg

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "sankhya_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sk_gpu_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(sg.GPU_LIB))
    names = declared_functions()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_csr_create_validates_before_touching_the_device():
    lib = sg.load_gpu()
    h = sg.GpuCsr()
    off = np.array([0, 3, 2], np.int32)
    col = np.array([0, 2, 1, 2], np.int32)
    val = np.array([1.0, 2.0, 3.0, 4.0])
    assert lib.sk_gpu_csr_create(C.byref(h), 2, 3, 4, sg.iptr(off), sg.iptr(col), sg.dptr(val)) == -1
    assert "row-offset" in sg.last_error()
    off = np.array([0, 2, 4], np.int32)
    bad = np.array([0, 3, 1, 2], np.int32)
    assert lib.sk_gpu_csr_create(C.byref(h), 2, 3, 4, sg.iptr(off), sg.iptr(bad), sg.dptr(val)) == -1
    assert "column index" in sg.last_error()


def test_device_csr_handle_has_the_reference_layout(tmp_path):
    """sk_gpu_csr is SankhyaCudaCSR field for field (three ints, then the
    three device pointers), so a caller's SankhyaCudaCSR passes as is."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include "sankhya_gpu.h"\n'
        "struct six { int rows, cols, nnz; void *ro, *ci, *v; };\n"
        "_Static_assert(sizeof(sk_gpu_csr) == sizeof(struct six), \"size\");\n"
        "_Static_assert(offsetof(sk_gpu_csr, nnz) == offsetof(struct six, nnz), \"nnz\");\n"
        "_Static_assert(offsetof(sk_gpu_csr, device_row_offsets) == offsetof(struct six, ro), \"ro\");\n"
        "_Static_assert(offsetof(sk_gpu_csr, device_values) == offsetof(struct six, v), \"v\");\n")
    import subprocess
    r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-I", str(ROOT / "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert C.sizeof(sg.GpuCsr) == 40


def test_destroy_of_a_foreign_handle_only_clears_it():
    """A handle the library did not create (here: made-up pointers) is
    cleared, and nothing is freed -- no device call, so this runs anywhere."""
    lib = sg.load_gpu()
    h = sg.GpuCsr(3, 4, 5, 0x1000, 0x2000, 0x3000)
    lib.sk_gpu_csr_destroy(C.byref(h))
    assert (h.rows, h.cols, h.nnz) == (0, 0, 0)
    assert h.device_row_offsets is None and h.device_values is None


def test_params_defaults_and_names():
    lib = sg.load_gpu()
    p = sg.PdlpParams()
    lib.sk_gpu_pdlp_params_default(C.byref(p))
    assert p.eps_optimal == 1e-6 and p.restart_period == 64 and p.power_iterations == 30
    assert p.adaptive_step == 1 and p.eta_min_factor == 1e-6 and p.eta_max_factor == 1e4
    assert lib.sk_gpu_result_name(0) == b"optimal" and lib.sk_gpu_status_name(3) == b"invalid_structure"
    o = sg.SkOptions()
    lib.sk_gpu_options_default(C.byref(o))
    assert o.primal_tol == 1e-7 and o.random_seed == 20260825 and o.threads == 1


def test_model_validate_matches_reference_rules():
    lib = sg.load_gpu()
    good = sg.Model.from_arrays(1, 1, [0, 1], [0], [1.0], [1.0], [1.0], [2.0], [0.0], [1.0])
    assert lib.sk_gpu_model_validate(good.ptr()) == 0
    bad = sg.Model.from_arrays(1, 1, [0, 1], [0], [np.inf], [1.0], [1.0], [2.0], [0.0], [1.0])
    assert lib.sk_gpu_model_validate(bad.ptr()) == sg.SK_ERR_STRUCTURE


def test_generators_are_deterministic_and_follow_their_laws():
    a = sg.Model.generated("skw_bounded_lp", 100, 200, 5, 42)
    b = sg.Model.generated("skw_bounded_lp", 100, 200, 5, 42)
    for u, v in zip(a.csc(), b.csc()):
        assert np.array_equal(u, v)
    p, i, x = a.csc()
    assert np.all(np.diff(p) == 5)
    for j in range(200):
        col = i[p[j]:p[j + 1]]
        assert np.all(np.diff(col) > 0)
    assert np.all(a.vec("clow") == 0) and np.all(a.vec("cupp") == 10)
    q = C.c_double(0)
    w = sg.load_workloads()
    z = sg.Model()
    assert w.skw_zipf_lp(C.byref(z.m), 20000, 40000, 1.2, 8.0, 2000, 0.3, 7, C.byref(q)) == 0
    z._generated = True
    deg = np.diff(z.csc()[0])
    assert -1.0 < q.value < 0.0
    assert abs(deg.mean() - 8.0) < 0.8 and deg.max() <= 2000 and deg.min() >= 1
    eq = z.vec("rlow") == z.vec("rupp")
    assert abs(eq.mean() - 0.3) < 0.02
    qp = sg.Model.generated("skw_sparse_qp", 60, 120, 3, 40, 3, 5)
    qp_, qi, qx = qp.q_csc()
    import scipy.sparse as sp
    Q = sp.csc_matrix((qx, qi, qp_), shape=(120, 120)).toarray()
    assert np.allclose(Q, Q.T) and np.linalg.eigvalsh(Q).min() > 0.0


def test_batch_validates_before_touching_the_device():
    """sk_gpu_batch_solve rejects what the persistent kernel cannot take
    before any device call: instances indexed by 16-bit local indices must
    stay at or below 65,536 rows and columns."""
    big = sg.Model.generated("skw_bounded_lp", 70000, 20000, 5, 3)
    small = sg.Model.generated("skw_bounded_lp", 30, 60, 5, 4)
    rc, _, _, _, _ = sg.batch_solve([small, big])
    assert rc == -1 and "65536" in sg.last_error()


class _Hcsr(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("nnz", C.c_int), ("p", C.POINTER(C.c_int)),
                ("j", C.POINTER(C.c_int)), ("v", C.POINTER(C.c_double))]


@pytest.mark.parametrize("m,n,nnz", [(300000, 200000, 1_500_000), (1000, 5, 1_200_000)])
def test_host_transpose_is_the_stable_counting_sort(m, n, nnz):
    """hcsr_transpose above 2^20 entries runs on host threads (each owns a
    range of output rows); its output must be the serial counting sort's:
    each output row in ascending source position, duplicates included, on
    a matrix with a few very long columns."""
    lib = sg.load_gpu()
    rng = np.random.default_rng(3)
    rows = np.sort(rng.integers(0, m, nnz))
    cols = rng.integers(0, n, nnz).astype(np.int32)
    vals = rng.standard_normal(nnz)
    rp = np.zeros(m + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    A = _Hcsr(m, n, nnz, sg.iptr(rp), sg.iptr(cols), sg.dptr(vals))
    T = _Hcsr()
    lib.hcsr_transpose.restype = C.c_int
    lib.hcsr_transpose.argtypes = [C.POINTER(_Hcsr), C.POINTER(_Hcsr)]
    lib.hcsr_free.restype = None
    lib.hcsr_free.argtypes = [C.POINTER(_Hcsr)]
    assert lib.hcsr_transpose(C.byref(A), C.byref(T)) == 0
    try:
        p = np.ctypeslib.as_array(T.p, shape=(n + 1,)).copy()
        j = np.ctypeslib.as_array(T.j, shape=(nnz,)).copy()
        v = np.ctypeslib.as_array(T.v, shape=(nnz,)).copy()
    finally:
        lib.hcsr_free(C.byref(T))
    order = np.argsort(cols, kind="stable")
    expect_p = np.concatenate([[0], np.cumsum(np.bincount(cols, minlength=n))])
    assert np.array_equal(p, expect_p)
    assert np.array_equal(j, rows[order]) and np.array_equal(v, vals[order])
