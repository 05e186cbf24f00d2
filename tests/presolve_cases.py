"""LP problems for the presolve tests: the hand-built cases of the
reference's tests/lp/test_presolve.cpp and a seeded random generator whose
models exercise every reduction (fixed, empty and free columns, empty,
singleton, forcing, redundant and ranged rows, infinite bounds, zero
coefficients, unsorted and duplicate entries)."""
from __future__ import annotations

import numpy as np

import sankhya_gpu as sg

INF = np.inf


def lp(rows, cols, triplets, obj, rhs, types, lower, upper, ranges=None):
    """An LpProblem assembled like the reference's tests: CSR from triplets
    (sorted, duplicates summed) and the default slack bounds of each row
    type (src/lp/LpProblem.cpp:7-28), narrowed by `ranges` {row: |r|}."""
    A = np.zeros((rows, cols))
    present = np.zeros((rows, cols), bool)
    for i, j, v in triplets:
        A[i, j] += v
        present[i, j] = True
    rp, cj, vv = [0], [], []
    for i in range(rows):
        for j in range(cols):
            if present[i, j]:
                cj.append(j)
                vv.append(A[i, j])
        rp.append(len(cj))
    sl = np.array([{"L": 0.0, "G": -INF, "E": 0.0}[t] for t in types]) if rows else np.zeros(0)
    su = np.array([{"L": INF, "G": 0.0, "E": 0.0}[t] for t in types]) if rows else np.zeros(0)
    for i, r in (ranges or {}).items():
        if types[i] == "L":
            su[i] = abs(r)
        elif types[i] == "G":
            sl[i] = -abs(r)
    return sg.LpProblem(rows, cols, np.array(rp), np.array(cj, np.int32), np.array(vv, np.float64),
                        np.array(obj, np.float64), np.array(rhs, np.float64), np.array(lower, np.float64),
                        np.array(upper, np.float64), sl, su, np.frombuffer("".join(types).encode(), np.uint8))


def tiny_lp():
    """min -x - y s.t. x + 2y <= 4, 3x + y <= 6, x, y >= 0: optimum
    (1.6, 1.2), -2.8 (reference tests/lp/test_presolve.cpp:28-40)."""
    return lp(2, 2, [(0, 0, 1.0), (0, 1, 2.0), (1, 0, 3.0), (1, 1, 1.0)], [-1.0, -1.0], [4.0, 6.0],
              "LL", [0.0, 0.0], [INF, INF])


def random_lp(rng: np.random.Generator) -> sg.LpProblem:
    """A small LP (m < 12, n < 14) whose rows mostly hold a point of the
    column box, so that presolve mostly reduces rather than rejects it."""
    m, n = int(rng.integers(0, 12)), int(rng.integers(1, 14))
    dens = rng.uniform(0.1, 0.6)
    rp, cj, vv = [0], [], []
    for _ in range(m):
        cols = np.flatnonzero(rng.random(n) < dens)
        if rng.random() < 0.15:
            cols = cols[:1]
        if rng.random() < 0.05:
            cols = cols[:0]
        vals = rng.choice([-3.0, -2.0, -1.0, 1.0, 2.0, 0.5, 3.0, 0.0], size=len(cols),
                          p=[.1, .1, .2, .25, .1, .1, .1, .05])
        if rng.random() < 0.1 and len(cols) > 1:
            cols = cols[::-1].copy()
        if rng.random() < 0.05 and len(cols) > 0:
            cols = np.append(cols, cols[0])
            vals = np.append(vals, 1.0)
        cj += list(cols)
        vv += list(vals)
        rp.append(len(cj))
    lo = rng.choice([0.0, -1.0, -INF, 2.0], size=n, p=[.5, .2, .2, .1])
    width = rng.choice([0.0, 1.0, 5.0, INF], size=n, p=[.1, .3, .3, .3])
    with np.errstate(invalid="ignore"):
        up = np.where(np.isfinite(lo), lo + width, np.where(np.isfinite(width), width, INF))
        x0 = np.where(np.isfinite(lo), lo, np.where(np.isfinite(up), up, 0.0))
        x0 = x0 + np.where(np.isfinite(up) & np.isfinite(lo), (up - lo) * 0.5 * rng.integers(0, 2, size=n), 0.0)
    A = np.zeros((m, n))
    for i in range(m):
        for k in range(rp[i], rp[i + 1]):
            A[i, cj[k]] += vv[k]
    if rng.random() < 0.8:
        rhs = A @ x0 + rng.choice([0.0, 0.0, 1.0, 2.0], size=m)
    else:
        rhs = rng.integers(-3, 8, size=m).astype(float)
    types = rng.choice(["L", "G", "E", "R"], size=m, p=[.4, .25, .2, .15])
    sl, su = np.zeros(m), np.zeros(m)
    for i, t in enumerate(types):
        if t == "L":
            su[i] = INF
        elif t == "G":
            sl[i] = -INF
        elif t == "R":
            su[i] = float(rng.integers(0, 4))
    obj = rng.choice([-1.0, 0.0, 1.0, 2.0], size=n, p=[.15, .25, .4, .2])
    return sg.LpProblem(m, n, np.array(rp), np.array(cj, np.int32), np.array(vv), obj, rhs, lo, up, sl, su,
                        np.array([ord("L" if t == "R" else t) for t in types], np.uint8))


def result_arrays(d):
    """A presolve result dict flattened to {key: ndarray} (for fixtures)."""
    out = {}
    for k, v in d.items():
        if k == "reduced":
            if v is not None:
                for kk, vv in v.arrays().items():
                    out[f"reduced/{kk}"] = np.asarray(vv)
        elif k == "reason":
            out[k] = np.array(v if v is not None else "")
        else:
            out[k] = np.asarray(v)
    return out


def mismatches(a, b):
    """Keys where two flattened results differ bit for bit."""
    bad = [k for k in sorted(set(a) | set(b)) if k not in a or k not in b]
    for k in sorted(set(a) & set(b)):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if x.shape != y.shape or x.dtype.kind != y.dtype.kind:
            bad.append(k)
        elif x.dtype.kind == "f":
            if x.tobytes() != y.astype(x.dtype).tobytes():
                bad.append(k)
        elif not np.array_equal(x, y):
            bad.append(k)
    return bad


def pack(cases):
    """Many flattened cases as one array per key plus per-case lengths
    (keys absent from a case get length -1)."""
    keys = sorted({k for c in cases for k in c})
    out = {}
    for k in keys:
        parts = [np.atleast_1d(np.asarray(c[k])) if k in c else None for c in cases]
        dt = next(p.dtype for p in parts if p is not None)
        data = np.concatenate([p for p in parts if p is not None])
        out[f"{k}#len"] = np.array([-1 if p is None else p.size for p in parts], np.int64)
        out[f"{k}#data"] = data if dt.kind == "U" else data.astype(dt)
        out[f"{k}#scalar"] = np.array([p is not None and np.asarray(c[k]).ndim == 0
                                       for p, c in zip(parts, cases)])
    return out


def unpack(z):
    """Inverse of pack: the list of flattened cases."""
    keys = sorted({name.split("#")[0] for name in z.files})
    count = len(z[f"{keys[0]}#len"])
    cases = [dict() for _ in range(count)]
    for k in keys:
        lens, data, scal = z[f"{k}#len"], z[f"{k}#data"], z[f"{k}#scalar"]
        at = 0
        for c, (ln, sc) in enumerate(zip(lens, scal)):
            if ln < 0:
                continue
            v = data[at:at + ln]
            cases[c][k] = v[0] if sc else v
            at += ln
    return cases
