"""MPS / QPS reader (host logic, no GPU).  Pinned three ways: the known
answers of the reference's own smoke tests (c/tests/mps_smoke.c,
c/tests/qps_smoke.c), the models the reference reader (sk_read_mps,
c/src/sk_core.c:318-378) produces for every fixture under tests/golden/mps/
(committed as tests/golden/mps_expected.npz by make_mps_golden.py), and,
where oracle/_ref is built, the reference reader run live on the same
files."""
from pathlib import Path

import numpy as np
import pytest

import sankhya_gpu as sg
import sk_oracle as so

GOLD = Path(__file__).resolve().parent / "golden"
FILES = sorted((GOLD / "mps").iterdir())
EXPECTED = np.load(GOLD / "mps_expected.npz")
FIELDS = ("nrow", "ncol", "objshift", "objsense", "Ap", "Ai", "Ax", "c", "rlow", "rupp", "clow", "cupp",
          "vartype", "rowname", "colname")


def _ours(path):
    lib = sg.load_gpu()
    m = sg.SkModel()
    st = lib.sk_gpu_read_mps(str(path).encode(), sg.C.byref(m))
    arrays = so.model_arrays(m) if st == 0 else None
    if st == 0:
        lib.sk_gpu_model_free(sg.C.byref(m))
    return st, arrays


def _same(a, b, name):
    for k in FIELDS + ("Qp", "Qi", "Qx"):
        if k in ("Qp", "Qi", "Qx") and k not in a and k not in b:
            continue
        assert k in a and k in b, (name, k)
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape and np.array_equal(x, y), (name, k, x, y)


@pytest.mark.parametrize("path", FILES, ids=[p.name for p in FILES])
def test_reader_matches_reference_fixtures(path):
    st, ours = _ours(path)
    assert st == int(EXPECTED[f"{path.name}/status"]), sg.last_error()
    if st == 0:
        exp = {k.split("/", 1)[1]: EXPECTED[k] for k in EXPECTED.files if k.startswith(path.name + "/")}
        _same(ours, exp, path.name)


@pytest.mark.parametrize("path", FILES, ids=[p.name for p in FILES])
def test_reader_matches_reference_live(path):
    got = so.ref_read_mps(path)
    if got is None:
        pytest.skip("oracle/_ref not built")
    st_ref, ref = got
    st, ours = _ours(path)
    assert st == st_ref
    if st == 0:
        _same(ours, ref, path.name)


def test_mps_smoke_known_answer(tmp_path):
    """c/tests/mps_smoke.c: 2 rows, 2 columns, 4 entries, c = (1, 2), X binary,
    the E row with RANGES -2 becomes [-1, 1]."""
    p = tmp_path / "smoke.mps"
    p.write_text("NAME TEST\nOBJSENSE\n MIN\nROWS\n N OBJ\n L CAP\n E BAL\nCOLUMNS\n X OBJ 1 CAP 1 BAL 1\n"
                 " Y OBJ 2 CAP 1 BAL -1\nRHS\n RHS1 OBJ 0 CAP 4 BAL 1\nRANGES\n RNG1 BAL -2\nBOUNDS\n BV BND X\n"
                 " LO BND Y 0\nENDATA\n")
    st, a = _ours(p)
    assert st == 0
    assert a["nrow"] == 2 and a["ncol"] == 2 and a["Ap"][-1] == 4
    assert a["c"].tolist() == [1.0, 2.0]
    assert a["vartype"].tolist() == [1, 0]           # SK_INTEGER, SK_CONTINUOUS
    assert a["rlow"][1] == -1.0 and a["rupp"][1] == 1.0
    assert a["rlow"][0] == -sg.SK_INFINITY and a["rupp"][0] == 4.0
    assert a["clow"].tolist() == [0.0, 0.0] and a["cupp"].tolist() == [1.0, sg.SK_INFINITY]


def test_qps_smoke_known_answer(tmp_path):
    """c/tests/qps_smoke.c: QUADOBJ lower triangle -> full symmetric Q,
    p = [0, 2, 4], i = [0, 1, 0, 1], x = [2, 1, 1, 4]."""
    p = tmp_path / "smoke.qps"
    p.write_text("NAME DIAGQP\nROWS\n N OBJ\n G R1\nCOLUMNS\n X1 OBJ -4 R1 1\n X2 OBJ 0\nRHS\n RHS1 R1 1\n"
                 "BOUNDS\n LO BND X1 0\nQUADOBJ\n X1 X1 2\n X1 X2 1\n X2 X2 4\nENDATA\n")
    st, a = _ours(p)
    assert st == 0
    assert a["Qp"].tolist() == [0, 2, 4] and a["Qi"].tolist() == [0, 1, 0, 1]
    assert a["Qx"].tolist() == [2.0, 1.0, 1.0, 4.0]


def test_reader_errors(tmp_path):
    lib = sg.load_gpu()
    m = sg.SkModel()
    assert lib.sk_gpu_read_mps(str(tmp_path / "missing.mps").encode(), sg.C.byref(m)) == sg.SK_ERR_IO
    bad = tmp_path / "badq.qps"
    bad.write_text("NAME B\nROWS\n N OBJ\n G R1\nCOLUMNS\n X OBJ 1 R1 1\nRHS\n RHS R1 1\nQUADOBJ\n X Z 1\nENDATA\n")
    assert lib.sk_gpu_read_mps(str(bad).encode(), sg.C.byref(m)) == sg.SK_ERR_STRUCTURE
    assert lib.sk_gpu_read_mps(None, sg.C.byref(m)) == sg.SK_ERR_ARG
