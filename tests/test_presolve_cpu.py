"""Presolve / postsolve / original-space gate on the host (no GPU needed).

Parity: sk_gpu_presolve must return the reference's PresolveResult bit for
bit (src/lp/Presolve.cpp:69-572) -- against fixtures the reference's own
presolve wrote (tests/golden/presolve, see make_presolve_golden.py) and,
where oracle/_ref is built, against the live reference on fresh seeded
models.  The hand-built cases restate tests/lp/test_presolve.cpp."""
import math
from pathlib import Path

import numpy as np
import pytest

import presolve_cases as pc
import sankhya_gpu as sg
import sk_oracle as so

INF = np.inf
GOLDEN = Path(__file__).resolve().parent / "golden" / "presolve"


def flat(problem, max_passes=20):
    return pc.result_arrays(sg.presolve(problem, max_passes))


def problem_from(case, prefix="in/"):
    g = {k[len(prefix):]: v for k, v in case.items() if k.startswith(prefix)}
    return sg.LpProblem(int(g["rows"]), int(g["cols"]), g["row_ptr"], g["col_idx"], g["values"], g["obj"],
                        g["rhs"], g["lower"], g["upper"], g["slack_lower"], g["slack_upper"], g["row_types"])


def expected_from(case):
    return {k[4:]: v for k, v in case.items() if k.startswith("out/")}


# -------------------------------------------------- reference unit tests

def test_singleton_row_becomes_a_bound():
    """tests/lp/test_presolve.cpp:49-80: x <= 3 as a row is absorbed into
    x's upper bound, relaxed outward by at most 1e-9 (1 + 3)."""
    p = pc.lp(2, 2, [(0, 0, 1.0), (1, 0, 1.0), (1, 1, 1.0)], [1.0, 1.0], [3.0, 10.0], "LE",
              [0.0, 0.0], [INF, INF])
    r = sg.presolve(p)
    assert r["status"] == sg.PRESOLVE_OK and r["removed_rows"] >= 1
    k = list(r["kept_columns"]).index(0)
    ub = r["reduced"].upper[k]
    assert 3.0 <= ub <= 3.0 + 1e-8
    assert ub == 3.0 + 1e-9 * 4.0      # exactly the documented relaxation


def test_redundant_row_dropped():
    """tests/lp/test_presolve.cpp:84-98: x, y in [0, 1], x + y <= 5."""
    p = pc.lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [-1.0, -1.0], [5.0], "L", [0.0, 0.0], [1.0, 1.0])
    r = sg.presolve(p)
    assert r["status"] == sg.PRESOLVE_OK and r["removed_rows"] == 1


def test_fixed_column_cascade_and_postsolve():
    """tests/lp/test_presolve.cpp:101-125: x fixed at 4 leaves the row as
    the singleton y = 6; no column survives; postsolve restores both."""
    p = pc.lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [0.0, 1.0], [10.0], "E", [4.0, 0.0], [4.0, INF])
    r = sg.presolve(p)
    assert r["status"] == sg.PRESOLVE_OK
    assert r["column_removed"][0] == 1 and abs(r["fixed_value"][0] - 4.0) < 1e-12
    assert r["removed_cols"] == 2 and r["reduced"].cols == 0
    x = sg.postsolve(p, np.full(len(r["kept_columns"]), 6.0))
    assert x.shape == (2,) and abs(x[0] - 4.0) < 1e-8 and abs(x[1] - 6.0) < 1e-8


def test_infeasible_bounds_detected():
    """tests/lp/test_presolve.cpp:128-141: x <= 3 and x >= 5."""
    p = pc.lp(2, 1, [(0, 0, 1.0), (1, 0, 1.0)], [0.0], [3.0, 5.0], "LG", [0.0], [INF])
    r = sg.presolve(p)
    assert r["status"] == sg.PRESOLVE_INFEASIBLE and r["reason"] is not None


def test_unbounded_empty_column_detected():
    """tests/lp/test_presolve.cpp:144-156."""
    p = pc.lp(1, 2, [(0, 0, 1.0)], [0.0, -1.0], [1.0], "L", [0.0, 0.0], [INF, INF])
    assert sg.presolve(p)["status"] == sg.PRESOLVE_UNBOUNDED


def test_crossed_bounds_name_the_culprit():
    p = pc.lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [1.0, 1.0], [1.0], "L", [0.0, 3.0], [1.0, 2.0])
    r = sg.presolve(p)
    assert (r["status"], r["reason"], r["reason_col"], r["reason_pass"]) == \
        (sg.PRESOLVE_INFEASIBLE, "column bounds crossed", 1, 0)
    q = pc.lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [1.0, 1.0], [1.0], "L", [0.0, 0.0], [INF, INF])
    q.slack_lower[0], q.slack_upper[0] = 2.0, 1.0          # rhs - 1 < A x < rhs - 2
    r = sg.presolve(q)
    assert (r["reason"], r["reason_row"]) == ("row bounds crossed", 0)


def test_forcing_row_pins_every_column():
    """x + y <= 0 with x, y >= 0: the minimum activity meets the upper
    bound, so both columns are pinned at their lower bounds."""
    p = pc.lp(2, 3, [(0, 0, 1.0), (0, 1, 1.0), (1, 1, 1.0), (1, 2, 1.0)], [1.0, 1.0, 1.0], [0.0, 4.0],
              "LL", [0.0, 0.0, 0.0], [INF, INF, INF])
    r = sg.presolve(p)
    assert r["status"] == sg.PRESOLVE_OK
    assert list(r["column_removed"][:2]) == [1, 1] and list(r["fixed_value"][:2]) == [0.0, 0.0]


def test_propagation_applies_lower_bounds_first():
    """Presolve.cpp:519-520: a row's proposed lower bounds are applied
    before its upper bounds, so when the two cross (a row whose bounds
    overlap within tolerance, lo = hi + 5e-8, and y nearly fixed) the
    lower one is taken and the upper one refused."""
    p = pc.lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [1.0, 1.0], [1.0], "L", [0.0, 0.0], [10.0, 1e-8])
    p.slack_upper[0] = -5e-8                         # lo = 1 + 5e-8 > hi = 1
    r = sg.presolve(p)
    assert r["status"] == sg.PRESOLVE_OK
    red = r["reduced"]
    lo = 1.0 - (-5e-8)
    rest_max = (1.0 * 10.0 + 1.0 * 1e-8) - 1.0 * 10.0     # activity max less x's own term
    limit = (lo - rest_max) / 1.0
    assert red.lower[0] == limit - 1e-9 * (1.0 + abs(limit))
    assert red.upper[0] == 10.0
    ref = so.ref_presolve(p)
    if ref is not None:
        assert not pc.mismatches(pc.result_arrays(r), pc.result_arrays(ref))


def test_emitted_rows_use_the_reference_encoding():
    """Kept rows re-encode lo <= A x <= hi as A x + s = rhs with rhs at the
    finite upper bound (Presolve.cpp:535-569)."""
    p = pc.lp(3, 3, [(0, 0, 1.0), (0, 1, 1.0), (1, 1, 1.0), (1, 2, -1.0), (2, 0, 1.0), (2, 2, 2.0)],
              [1.0, 1.0, 1.0], [5.0, 1.0, 2.0], "LGL", [0.0, 0.0, 0.0], [10.0, 10.0, 10.0],
              ranges={0: 3.0})
    r = sg.presolve(p)
    red = r["reduced"]
    for k, i in enumerate(r["kept_rows"]):
        lo = p.rhs[i] - p.slack_upper[i]
        hi = p.rhs[i] - p.slack_lower[i]
        if math.isfinite(hi):
            assert red.rhs[k] == hi and red.slack_lower[k] == 0.0
            assert red.slack_upper[k] == (hi - lo if math.isfinite(lo) else INF)
            assert chr(red.row_types[k]) == ("E" if math.isfinite(lo) and hi - lo <= 1e-11 else "L")
        else:
            assert (red.rhs[k], red.slack_lower[k], red.slack_upper[k], chr(red.row_types[k])) == \
                (lo, -INF, 0.0, "G")


def test_postsolve_places_kept_and_fixed_values():
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = pc.random_lp(rng)
        r = sg.presolve(p)
        if r["status"] != sg.PRESOLVE_OK:
            continue
        rx = rng.normal(size=len(r["kept_columns"]))
        x = sg.postsolve(p, rx)
        assert np.array_equal(x[r["kept_columns"]], rx)
        gone = r["column_removed"].astype(bool)
        assert np.array_equal(x[gone], r["fixed_value"][gone])
        assert not np.any(gone[r["kept_columns"]])
        # a short reduced vector leaves the missing kept columns at 0
        x2 = sg.postsolve(p, rx[:len(rx) // 2])
        assert np.all(x2[r["kept_columns"][len(rx) // 2:]] == 0.0)


def test_primal_residual_closed_form():
    """src/lp/LpSolver.cpp:23-50: max(row violation / (1 + ||rhs||_inf),
    bound violation)."""
    p = pc.lp(2, 2, [(0, 0, 1.0), (0, 1, 2.0), (1, 0, 3.0), (1, 1, 1.0)], [-1.0, -1.0], [4.0, 6.0], "LL",
              [0.0, 0.0], [INF, INF])
    assert sg.primal_residual(p, [1.6, 1.2]) <= 1e-15
    # row 1: 3*2 + 1 = 7 > 6 by 1 -> 1 / (1 + 6); no bound violated
    assert sg.primal_residual(p, [2.0, 1.0]) == 1.0 / 7.0
    # x below its lower bound by 0.5 dominates a satisfied row set
    assert sg.primal_residual(p, [-0.5, 0.0]) == 0.5


def test_problem_from_model_encoding():
    """sk_model rows map to the reference encoding; the 1e30 sentinel
    becomes IEEE infinity."""
    big = sg.SK_INFINITY
    m = sg.Model.from_arrays(3, 2, [0, 2, 4], [0, 1, 1, 2], [1.0, 2.0, 3.0, 4.0], [1.0, -1.0],
                             [-big, 1.0, -2.0], [5.0, big, 2.0], [0.0, -big], [big, 3.0])
    p = sg.LpProblem.from_model(m)
    assert list(p.rhs) == [5.0, 1.0, 2.0]
    assert list(p.slack_lower) == [0.0, -INF, 0.0] and list(p.slack_upper) == [INF, 0.0, 4.0]
    assert bytes(p.row_types).decode() == "LGL"
    assert list(p.lower) == [0.0, -INF] and list(p.upper) == [INF, 3.0]
    assert list(p.row_ptr) == [0, 1, 3, 4] and list(p.col_idx) == [0, 0, 1, 1]
    assert list(p.values) == [1.0, 2.0, 3.0, 4.0]


def test_invalid_problem_rejected():
    p = pc.tiny_lp()
    p.col_idx[0] = 7
    with pytest.raises(RuntimeError, match="column index"):
        sg.presolve(p)


# -------------------------------------------------- fixtures of the reference

@pytest.mark.parametrize("path", sorted(GOLDEN.glob("*.npz")), ids=lambda p: p.stem)
def test_matches_reference_fixtures(path):
    z = np.load(path)
    cases = pc.unpack(z) if path.stem == "random" else [dict(z)]
    statuses = set()
    for k, case in enumerate(cases):
        got = flat(problem_from(case))
        bad = pc.mismatches(got, expected_from(case))
        assert not bad, f"case {k}: {bad}"
        statuses.add(int(got["status"]))
    if path.stem == "random":   # the sample covers every outcome
        assert statuses == {0, 1, 2}


# -------------------------------------------------- the live reference

ref_missing = pytest.mark.skipif(so.load_presolve_ref() is None, reason="oracle/_ref not built")


@ref_missing
def test_matches_live_reference_on_seeded_models():
    rng = np.random.default_rng(777)
    reductions = 0
    for k in range(1500):
        p = pc.random_lp(rng)
        got = flat(p)
        bad = pc.mismatches(got, pc.result_arrays(so.ref_presolve(p)))
        assert not bad, f"model {k}: {bad}"
        reductions += int(got["status"]) == 0 and (got["removed_rows"] + got["removed_cols"]) > 0
    assert reductions > 300


@ref_missing
def test_pass_cap_matches_reference():
    rng = np.random.default_rng(31)
    for k in range(300):
        p = pc.random_lp(rng)
        for passes in (1, 2):
            bad = pc.mismatches(flat(p, passes), pc.result_arrays(so.ref_presolve(p, passes)))
            assert not bad, f"model {k} passes {passes}: {bad}"


@ref_missing
def test_postsolve_matches_reference():
    rng = np.random.default_rng(12)
    for _ in range(200):
        p = pc.random_lp(rng)
        r = sg.presolve(p)
        if r["status"] != sg.PRESOLVE_OK:
            continue
        rx = rng.normal(size=len(r["kept_columns"]))
        assert sg.postsolve(p, rx).tobytes() == so.ref_postsolve(p, rx).tobytes()


# -------------------------------------------------- pipeline paths that never reach the device

def test_solve_lp_ends_at_presolve_verdicts():
    """LpSolver.cpp:138-145: INFEASIBLE / UNBOUNDED from presolve end the
    solve with no solution."""
    inf = pc.lp(2, 1, [(0, 0, 1.0), (1, 0, 1.0)], [0.0], [3.0, 5.0], "LG", [0.0], [INF])
    unb = pc.lp(1, 2, [(0, 0, 1.0)], [0.0, -1.0], [1.0], "L", [0.0, 0.0], [INF, INF])
    for p, status in ((inf, sg.LP_INFEASIBLE), (unb, sg.LP_UNBOUNDED)):
        sol, _ = sg.solve_lp(p)
        assert sol.status == status and not sol.has_x and not sol.used_first_order


def test_solve_lp_fully_presolved_model():
    """LpSolver.cpp:157-164: nothing left to optimise is OPTIMAL once the
    postsolved point passes the original-space gate."""
    p = pc.lp(1, 2, [(0, 0, 1.0), (0, 1, 1.0)], [0.0, 1.0], [10.0], "E", [4.0, 0.0], [4.0, INF])
    sol, x = sg.solve_lp(p)
    assert sol.status == sg.LP_OPTIMAL and sol.has_x and not sol.used_first_order
    assert (sol.presolve_removed_rows, sol.presolve_removed_cols) == (1, 2)
    assert x[0] == 4.0 and abs(x[1] - 6.0) < 1e-8 and sol.objective_value == x[1]
    assert sol.primal_residual == sg.primal_residual(p, x) <= 1e-6


def test_solve_lp_rejects_invalid_input():
    p = pc.tiny_lp()
    p.row_ptr[1] = 5
    with pytest.raises(RuntimeError):
        sg.solve_lp(p)


@ref_missing
def test_matches_live_reference_on_a_power_law_lp():
    """The config-1 generator at a fifth of its size (200,000 x 400,000,
    about 3.2M nonzeros): every pass runs and thousands of rows and columns
    go, bit for bit as the reference's presolve."""
    import ctypes as C
    w = sg.load_workloads()
    model = sg.Model()
    q = C.c_double()
    assert w.skw_zipf_lp(C.byref(model.m), 200000, 400000, 1.2, 8.0, 2000, 0.3, 20260826, C.byref(q)) == 0
    model._generated = True
    p = sg.LpProblem.from_model(model)
    got = flat(p)
    assert int(got["removed_rows"]) > 1000 and int(got["removed_cols"]) > 1000
    bad = pc.mismatches(got, pc.result_arrays(so.ref_presolve(p)))
    assert not bad, bad
