"""Row-tile planner of the CSR-stream products (host logic, no GPU): tiles
cover every row once, in order; a tile holds at most the entry cap (2048
for a large matrix) and 1024 rows unless it is a single longer row; an
empty matrix still gets one tile.  The cap halves, down to 256, while the
matrix would have fewer than 1024 tiles."""
import ctypes as C

import numpy as np
import pytest

import sankhya_gpu as sg

TILE_NNZ, TILE_ROWS = 2048, 1024


def _plan(lengths, tile_nnz=TILE_NNZ):
    lib = sg.load_gpu()
    rp = np.zeros(len(lengths) + 1, dtype=np.int32)
    rp[1:] = np.cumsum(lengths)
    out = C.POINTER(C.c_int)()
    f = lib.skg_plan_tiles
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_int))]
    nt = f(rp.ctypes.data_as(C.POINTER(C.c_int)), len(lengths), tile_nnz, C.byref(out))
    assert nt >= 1
    tiles = np.array([out[k] for k in range(nt + 1)])
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    libc.free(C.cast(out, C.c_void_p))
    return rp, tiles


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tiles_partition_rows(seed):
    rng = np.random.default_rng(seed)
    lengths = np.minimum(rng.zipf(1.3, size=20000), 6000)
    lengths[rng.integers(0, 20000, 500)] = 0
    rp, t = _plan(lengths)
    assert t[0] == 0 and t[-1] == len(lengths) and np.all(np.diff(t) > 0)
    for a, b in zip(t[:-1], t[1:]):
        nnz = rp[b] - rp[a]
        if nnz > TILE_NNZ:
            assert b - a == 1
        else:
            assert b - a <= TILE_ROWS
            # greedy: the next row would not have fitted
            if b < len(lengths) and b - a < TILE_ROWS and lengths[b] <= TILE_NNZ:
                assert nnz + lengths[b] > TILE_NNZ


def test_tiles_edge_cases():
    _, t = _plan(np.zeros(0, dtype=np.int64))
    assert t.tolist() == [0, 0]
    _, t = _plan(np.zeros(3000, dtype=np.int64))        # empty rows: row cap only
    assert t.tolist() == [0, 1024, 2048, 3000]
    _, t = _plan(np.array([2048, 2049, 1, 2047]))
    assert t.tolist() == [0, 1, 2, 4]


def test_tiles_smaller_cap():
    rp, t = _plan(np.full(100, 300), tile_nnz=512)     # one 300-entry row per tile
    assert np.all(np.diff(t) == 1)
    rp, t = _plan(np.full(100, 100), tile_nnz=512)     # five rows of 100 per tile
    assert np.all(np.diff(t)[:-1] == 5)


def test_tile_cap_rule():
    f = sg.load_gpu().skg_tile_nnz
    f.restype, f.argtypes = C.c_int, [C.c_longlong]
    assert f(10_000_000) == 2048 and f(2048 * 1024) == 2048
    assert f(2048 * 1024 - 1) == 1024 and f(1024 * 1024) == 1024
    assert f(600 * 1024) == 512 and f(300 * 1024) == 256 and f(1000) == 256 and f(0) == 256
