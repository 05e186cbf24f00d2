#!/usr/bin/env python3
"""Benchmark: restarted PDHG (PDLP) on one MI355X.

Workload (the headline configuration): bounded LP with m = 1,000,000 rows and
n = 2,000,000 columns, Zipf(1.2) column degrees with mean 8 capped at 2000,
log-uniform |a_ij| in [1e-3, 1e3], 30% equality rows, bounds [0, 10]
(generator skw_zipf_lp, seed 20260826), synthetic and generated in-process.

A "step" is one PDHG iteration of the adaptive-step loop (dual prox, A^T y
with fused primal prox, A x with fused step decision).  `value` is the mixed
engine's whole-solve rate: a cold-start solve of exactly --steps iterations
through the real loop -- restart windows, FP64 evaluations and read-backs,
restarts, re-anchoring of the FP32 correction, export of the result -- in
PDHG iterations per second; the FP64 engine ("plain FP64 GPU PDHG") is timed
the same way beside it, and the raw inner loop of both is reported too.
Unless --skip-solve, both engines also solve the LP from a cold start toward
a relative KKT of 1e-6 (time limit --solve-time-limit) and report the times
at which they first reach 1e-3 ... 1e-6.  The roofline object reports the
dominant kernel's algorithmic HBM bytes divided by its HIP-event-timed mean
duration on the library stream.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-solve] [--skip-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "the-solver_amd"))
import sankhya_gpu as sg  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "time-to-1e-6 KKT & PDHG iters/s vs plain FP64 GPU PDHG; SpMV HBM GB/s"
CFG = dict(m=1_000_000, n=2_000_000, zipf_s=1.2, mean_degree=8.0, max_degree=2000,
           eq_fraction=0.3, seed=20260826)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(quick=False):
    w = sg.load_workloads()
    model = sg.Model()
    q = C.c_double()
    m, n = (CFG["m"], CFG["n"]) if not quick else (20000, 40000)
    rc = w.skw_zipf_lp(C.byref(model.m), m, n, CFG["zipf_s"], CFG["mean_degree"],
                       CFG["max_degree"], CFG["eq_fraction"], CFG["seed"], C.byref(q))
    if rc != 0:
        raise RuntimeError("workload generation failed")
    model._generated = True
    return model, q.value


class Session:
    def __init__(self, lib, model, arrays, **params):
        self.lib = lib
        self.model = model
        rp, cj, v, vec = arrays
        self.P = sg.default_pdlp_params(**params)
        self.h = C.c_void_p()
        rc = lib.sk_gpu_pdlp_session_create(model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj),
                                            sg.dptr(v), sg.dptr(vec["c"]), sg.dptr(vec["rlow"]),
                                            sg.dptr(vec["rupp"]), sg.dptr(vec["clow"]), sg.dptr(vec["cupp"]),
                                            C.byref(self.P), C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"session create failed: {sg.last_error()}")

    def iterate(self, k, timed=False):
        ms = np.zeros(5)
        rc = self.lib.sk_gpu_pdlp_session_iterate(self.h, k, sg.dptr(ms) if timed else None)
        if rc != 0:
            raise RuntimeError(f"iterate failed: {sg.last_error()}")
        return ms

    def solve(self, x0=None, y0=None):
        x = np.zeros(self.model.ncol)
        y = np.zeros(self.model.nrow)
        st = sg.PdlpStats()
        x0 = np.ascontiguousarray(x0, np.float64) if x0 is not None else None
        y0 = np.ascontiguousarray(y0, np.float64) if y0 is not None else None
        rc = self.lib.sk_gpu_pdlp_session_solve(self.h, sg.dptr(x0) if x0 is not None else None,
                                                sg.dptr(y0) if y0 is not None else None,
                                                sg.dptr(x), sg.dptr(y), C.byref(st))
        return rc, x, y, st

    def update(self, c, rl, ru, l, u):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (c, rl, ru, l, u)]
        if self.lib.sk_gpu_pdlp_session_update(self.h, *[sg.dptr(a) for a in arrs]) != 0:
            raise RuntimeError(sg.last_error())

    def set_limits(self, max_iterations, time_limit, eps):
        if self.lib.sk_gpu_pdlp_session_set_limits(self.h, int(max_iterations), float(time_limit), float(eps)) != 0:
            raise RuntimeError(sg.last_error())

    def close(self):
        if self.h:
            self.lib.sk_gpu_pdlp_session_destroy(self.h)
            self.h = C.c_void_p()


def pmc_traffic(name="spmv_lp_1Mx2M_pmc.md"):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes (calibrated against a copy kernel), from the committed summary."""
    path = ROOT / "profiles" / name
    out = {}
    if path.exists():
        block = path.read_text().split("```traffic", 1)
        if len(block) == 2:
            for line in block[1].split("```", 1)[0].strip().splitlines():
                k, v = line.split()
                out[k] = int(v)
    return out


def kernel_bytes(m, n, nnz, s, s_v, accept_rate):
    """Algorithmic HBM bytes of the three launches of one adaptive iteration
    (s = iterate element size, s_v = matrix value size, int32 indices)."""
    k1 = s * (5 * m + m) + accept_rate * s * 2 * m           # y, Ax, Ax_prev, rl, ru -> y_cand (+ y_sum)
    k2 = nnz * (s_v + 4) + 4 * (n + 1) + s * m \
        + s * (4 * n + n) + accept_rate * s * 2 * n          # A^T y_cand; x, c, l, u -> x_cand (+ x_sum)
    k3 = nnz * (s_v + 4) + 4 * (m + 1) + s * n + s * (3 * m + m)  # A x_cand; y, y_cand, Ax -> Ax_cand
    return k1, k2, k3


def cpu_baseline(model, iters):
    """Reference C core (oracle/_ref, compiled from the reference sources) or,
    where it is absent, the oracle's restatement: the CPU PDHG of
    sk_solve(lp_engine = FIRST_ORDER).  The rate is the difference of two
    iteration budgets (iters and iters / 5), so sk_solve's own set-up does not
    count; one thread, and all the host's threads for the reference (its
    A^T y loop is OpenMP-parallel, c/src/sk_solve.c:40-50)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import sk_oracle as so

    kind = "reference" if so.load_ref() is not None else "port"

    def rate(threads):
        out = []
        for budget in (max(iters // 5, 1), iters):
            t0 = time.perf_counter()
            rc, s = so.first_order_ref(model, budget, threads) if kind == "reference" else \
                so.first_order_orc(model, budget)
            out.append((int(s.iterations), time.perf_counter() - t0))
        (i0, t0), (i1, t1) = out
        if i1 <= i0 or t1 - t0 < 0.2 * t1:      # set-up not resolvable: whole-run rate
            return i1 / max(t1, 1e-9), i1, t1
        return (i1 - i0) / (t1 - t0), i1, t1

    v1, done, dt = rate(1)
    line = {"value": v1, "unit": "PDHG iterations/s", "cores": 1, "kind": kind,
            "sample": f"{done} and {max(iters // 5, 1)} iterations of the reference CPU PDHG (sk_solve, "
                      f"lp_engine=FIRST_ORDER, threads=1) on the same LP; rate = difference of the two "
                      f"budgets ({dt:.1f} s for the larger, set-up excluded)"}
    if kind == "reference":
        nproc = min(os.cpu_count() or 1, 16)
        vn, _, _ = rate(nproc)
        line["threads_all"] = {"value": vn, "cores": nproc}
    return line


def _solve_budget(S, iters, eps=1e-12):
    """One cold-start solve of exactly `iters` iterations through the real
    loop (windows, FP64 evaluations and read-backs, restarts, re-anchoring,
    export of the result to the host), wall-clocked between two device
    synchronisations."""
    S.set_limits(iters, 0.0, eps)
    sg.device_synchronize()
    t = time.perf_counter()
    rc, x, y, st = S.solve()
    sg.device_synchronize()
    wall = time.perf_counter() - t
    if rc != 0 or st.iterations != iters:
        raise RuntimeError(f"budget solve: rc={rc} iterations={st.iterations} ({sg.last_error()})")
    return wall, st


def _tiling_params(args):
    """Graph replay and the products' tiling of the headline sessions: the
    autotuner by default; --tiling T,G,S fixes tile cap, grid cap and
    streaming (0 rule, 1 off, 2 on) -- profiles use it to run the tiling the
    tuner picked without timing candidates inside the profiled process."""
    kw = dict(use_graph=0 if args.no_graph else 1)
    if args.tiling:
        v = [int(x) for x in args.tiling.split(",")] + [0, 0]
        kw.update(autotune=0, tile_nnz=v[0], spmv_grid=v[1], stream_loads=v[2], probe_grid=v[3],
                  xcache_bytes=v[4])
    else:
        kw.update(autotune=0 if args.no_tune else 1)
    return kw


def workload_lp1(args):
    """Headline: the 1M x 2M power-law LP.  `value` is the mixed engine's
    whole-solve rate: --steps iterations from a cold start through the real
    loop, everything it does inside the timed region."""
    lib = sg.load_gpu()
    t0 = time.perf_counter()
    model, q = make_workload(args.quick)
    p, i, x = model.csc()
    rp, cj, v = sg.csc_to_csr(model.nrow, model.ncol, p, i, x)
    vec = {k: model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")}
    arrays = (rp, cj, v, vec)
    m, n, nnz = model.nrow, model.ncol, int(rp[-1])
    log(f"workload m={m} n={n} nnz={nnz} zipf_q={q:.6f} built in {time.perf_counter() - t0:.1f}s")
    _warm_device(lib, 1)

    results = {}
    for name, prec in (("mixed", sg.PREC_MIXED), ("fp64", sg.PREC_FP64)):
        t = time.perf_counter()
        S = Session(lib, model, arrays, precision=prec, **_tiling_params(args))
        setup = time.perf_counter() - t
        _solve_budget(S, max(args.warmup, 1))                 # graph capture, caches
        wall, st = _solve_budget(S, args.steps)
        r = dict(setup_s=setup, iters_per_s=args.steps / wall, ms_per_iter=1e3 * wall / args.steps,
                 budget=dict(iterations=st.iterations, fp32_iterations=st.fp32_iterations,
                             restarts=st.restarts, reanchors=st.reanchors, host_syncs=st.host_syncs,
                             kernel_launches=st.kernel_launches),
                 tiling=dict(tuned=st.tuned, tile_nnz=st.tile_nnz, spmv_grid=st.spmv_grid,
                             probe_grid=st.probe_grid, xcache_bytes=st.xcache_bytes,
                             stream_loads=st.stream_loads, tune_s=round(st.tune_seconds, 3)))
        # the inner loop alone: raw iterations, no evaluation or restart
        S.iterate(args.warmup)
        t = time.perf_counter()
        S.iterate(args.steps)
        r["inner_iters_per_s"] = args.steps / (time.perf_counter() - t)
        ms = S.iterate(min(args.steps, 200), timed=True)
        r["kernel_ms"] = ms.tolist()
        log(f"{name}: whole solve {r['iters_per_s']:.1f} it/s, inner loop {r['inner_iters_per_s']:.1f} it/s, "
            f"kernels(ms) {np.round(ms, 4).tolist()}")
        if not args.skip_solve:
            S.set_limits(500000, args.solve_time_limit, 1e-6)
            t = time.perf_counter()
            rc, xs, ys, st = S.solve()
            solve_wall = time.perf_counter() - t
            kkt = np.zeros(5)
            lib.sk_gpu_relative_kkt(m, n, nnz, sg.iptr(rp), sg.iptr(cj), sg.dptr(v), sg.dptr(vec["c"]),
                                    sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]), sg.dptr(vec["clow"]),
                                    sg.dptr(vec["cupp"]), sg.dptr(xs), sg.dptr(ys), sg.dptr(kkt))
            d = st.as_dict()
            r.update(solve_rc=rc, converged=bool(st.converged), iterations=st.iterations,
                     fp32_iterations=st.fp32_iterations, fp64_iterations=st.fp64_iterations,
                     loop_s=st.loop_seconds, solve_s=solve_wall, end_to_end_s=setup + solve_wall,
                     restarts=st.restarts, reanchors=st.reanchors,
                     time_to_merit_s=dict(zip(("1e-3", "1e-4", "1e-5", "1e-6"),
                                              [v_ if v_ > 0 else None for v_ in d["seconds_to_merit"]])),
                     iterations_to_merit=dict(zip(("1e-3", "1e-4", "1e-5", "1e-6"),
                                                  [v_ if v_ > 0 else None for v_ in d["iterations_to_merit"]])),
                     original_kkt=kkt[:3].tolist(), objective=float(vec["c"] @ xs))
            log(f"{name} solve: rc={rc} conv={st.converged} it={st.iterations} loop={st.loop_seconds:.2f}s "
                f"kkt={kkt[:3]} to-merit {d['seconds_to_merit']}")
        S.close()
        results[name] = r

    if not args.skip_solve and not args.skip_ndg:
        # the same mixed engine with normalized-duality-gap restarts
        S = Session(lib, model, arrays, precision=sg.PREC_MIXED, restart_rule=sg.RESTART_NDG, **_tiling_params(args))
        S.set_limits(500000, args.solve_time_limit, 1e-6)
        rc, xs, ys, st = S.solve()
        d = st.as_dict()
        kkt = np.zeros(5)
        lib.sk_gpu_relative_kkt(m, n, nnz, sg.iptr(rp), sg.iptr(cj), sg.dptr(v), sg.dptr(vec["c"]),
                                sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]), sg.dptr(vec["clow"]),
                                sg.dptr(vec["cupp"]), sg.dptr(xs), sg.dptr(ys), sg.dptr(kkt))
        results["mixed_ndg"] = dict(converged=bool(st.converged), iterations=st.iterations, loop_s=st.loop_seconds,
                                    time_to_merit_s=dict(zip(("1e-3", "1e-4", "1e-5", "1e-6"),
                                                             [v_ if v_ > 0 else None for v_ in d["seconds_to_merit"]])),
                                    original_kkt=kkt[:3].tolist(), restarts=st.restarts)
        log(f"mixed NDG solve: conv={st.converged} it={st.iterations} loop={st.loop_seconds:.2f}s kkt={kkt[:3]} "
            f"to-merit {d['seconds_to_merit']}")
        S.close()

    mix, f64 = results["mixed"], results["fp64"]
    k1, k2, k3 = kernel_bytes(m, n, nnz, 4, 4, 1.0)
    kernels = {"dual_prox": (k1, mix["kernel_ms"][0]), "AT_spmv_primal": (k2, mix["kernel_ms"][1]),
               "A_spmv_probe": (k3, mix["kernel_ms"][2])}
    # [3] is the one-workgroup accept/reject decision (no HBM stream), [4] the iteration
    dom = max(kernels, key=lambda k: kernels[k][1])
    bytes_dom, ms_dom = kernels[dom]
    achieved = bytes_dom / (ms_dom * 1e-3) / 1e9
    spmv_gbs = {k: kernels[k][0] / (kernels[k][1] * 1e-3) / 1e9 for k in kernels}
    # the products gather one 4-byte vector element per entry at random
    # positions: their second ceiling is the device's random-gather rate
    # for a table of the gathered vector's size
    gather = {}
    for name, table in (("AT_spmv_primal", m), ("A_spmv_probe", n)):
        rate = C.c_double()
        if lib.sk_gpu_gather_rate(table, nnz, 10, C.byref(rate)) == 0 and rate.value > 0:
            got = nnz / (kernels[name][1] * 1e-3)
            gather[name] = {"gathers": nnz, "table_bytes": 4 * table, "achieved_G_per_s": round(got / 1e9, 1),
                            "probe_G_per_s": round(rate.value / 1e9, 1), "frac": round(got / rate.value, 3)}
    line = {
        "metric": METRIC,
        "value": round(mix["iters_per_s"], 2),
        "unit": "PDHG iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mix["ms_per_iter"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded generator, in-process)",
        "config": {"workload": f"bounded LP m={m} n={n} nnz={nnz}, Zipf(1.2) column degrees mean 8 cap 2000, "
                               f"|a| log-uniform [1e-3,1e3], 30% equality rows, bounds [0,10]",
                   "engine": "adaptive-step restarted PDHG, FP32 iterates anchored to FP64, FP64 checks",
                   "step": "one PDHG iteration; value = whole cold-start solve of `steps` iterations "
                           "(evaluations, restarts, re-anchoring and result export included)"},
        "fp64_iters_per_s": round(f64["iters_per_s"], 2),
        "fp64_ms_per_step": round(f64["ms_per_iter"], 4),
        "speedup_iters_vs_fp64": round(mix["iters_per_s"] / f64["iters_per_s"], 3),
        "inner_loop_iters_per_s": {"mixed": round(mix["inner_iters_per_s"], 1),
                                   "fp64": round(f64["inner_iters_per_s"], 1)},
        "budget_solve": {"mixed": mix["budget"], "fp64": f64["budget"]},
        "tiling": {"mixed": mix["tiling"], "fp64": f64["tiling"]},
        "setup_s": {"mixed": round(mix["setup_s"], 3), "fp64": round(f64["setup_s"], 3)},
        "spmv_gbs": {k: round(val, 1) for k, val in spmv_gbs.items()},
        "gather_roofline": gather,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic().get(dom) if not args.quick else None,
                     "algorithmic_bytes": int(bytes_dom), "mean_ms": round(ms_dom, 5)},
    }
    if not args.skip_solve:
        keys = ("converged", "iterations", "fp32_iterations", "fp64_iterations", "loop_s", "end_to_end_s",
                "time_to_merit_s", "iterations_to_merit", "original_kkt", "objective", "restarts", "reanchors")
        line["solve"] = {"mixed": {k: mix[k] for k in keys}, "fp64": {k: f64[k] for k in keys}}
        line["time_to_kkt_1e-6_s"] = {"mixed": mix["time_to_merit_s"]["1e-6"], "fp64": f64["time_to_merit_s"]["1e-6"]}
        line["time_to_kkt_1e-4_s"] = {"mixed": mix["time_to_merit_s"]["1e-4"], "fp64": f64["time_to_merit_s"]["1e-4"]}
        for tol in ("1e-4", "1e-6"):
            a, b = mix["time_to_merit_s"][tol], f64["time_to_merit_s"][tol]
            if a and b:
                line[f"speedup_time_to_kkt_{tol}_vs_fp64"] = round(b / a, 3)
        if "mixed_ndg" in results:
            line["solve"]["mixed_ndg_restarts"] = results["mixed_ndg"]
        # error of the mixed result against the FP64 engine's, both in the
        # original space, FP64 on the device
        line["mixed_vs_fp64"] = {"objective_rel_diff": abs(mix["objective"] - f64["objective"]) /
                                 (1.0 + abs(f64["objective"])),
                                 "original_kkt_mixed": mix["original_kkt"], "original_kkt_fp64": f64["original_kkt"]}
    if not args.skip_cpu:
        log("CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(model, args.cpu_iters)
    return line


# ------------------------------------------------------------ other configs

def _arrays(model):
    p, i, x = model.csc()
    rp, cj, v = sg.csc_to_csr(model.nrow, model.ncol, p, i, x)
    vec = {k: model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")}
    return rp, cj, v, vec


def _session_qp(lib, model, arrays, q, **params):
    """A PDLP session with an optional full symmetric Hessian q = (p, i, x)."""
    rp, cj, v, vec = arrays
    S = Session.__new__(Session)
    S.lib, S.model = lib, model
    S.P = sg.default_pdlp_params(**params)
    S.h = C.c_void_p()
    qa = [None, None, None, 0] if q is None else [sg.iptr(q[0]), sg.iptr(q[1]), sg.dptr(q[2]), int(q[0][-1])]
    rc = lib.sk_gpu_pdlp_session_create_qp(model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj),
                                           sg.dptr(v), *qa, sg.dptr(vec["c"]), sg.dptr(vec["rlow"]),
                                           sg.dptr(vec["rupp"]), sg.dptr(vec["clow"]), sg.dptr(vec["cupp"]),
                                           C.byref(S.P), C.byref(S.h))
    if rc != 0:
        raise RuntimeError(f"session create failed: {sg.last_error()}")
    return S


def _kkt(lib, model, arrays, q, x, y):
    rp, cj, v, vec = arrays
    out = np.zeros(5)
    qa = [None, None, None, 0] if q is None else [sg.iptr(q[0]), sg.iptr(q[1]), sg.dptr(q[2]), int(q[0][-1])]
    rc = lib.sk_gpu_relative_kkt_qp(model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj), sg.dptr(v),
                                    *qa, sg.dptr(vec["c"]), sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]),
                                    sg.dptr(vec["clow"]), sg.dptr(vec["cupp"]), sg.dptr(x), sg.dptr(y),
                                    sg.dptr(out))
    if rc != 0:
        raise RuntimeError(sg.last_error())
    return out


def _warm_device(lib, use_graph):
    """One tiny session per precision first, so that the HIP runtime's
    one-time start-up (device initialisation, code-object loading) is not
    charged to the set-up of whichever engine runs first."""
    tiny = sg.Model.generated("skw_bounded_lp", 4, 8, 2, 1)
    tarr = _arrays(tiny)
    for prec in (sg.PREC_MIXED, sg.PREC_FP64):
        S = _session_qp(lib, tiny, tarr, None, precision=prec, use_graph=use_graph)
        S.iterate(2)
        S.solve()
        S.close()


def _engines(lib, model, arrays, q, args, use_graph=0, repeats=3):
    """Iterations/s and cold time to 1e-6 for the mixed and the FP64 engine:
    the median of `repeats` cold solves of the resident session (the loop on
    the device) and, end to end, the session set-up (scaling, permutation,
    upload, ||A||) plus one solve."""
    _warm_device(lib, use_graph)
    res = {}
    for name, prec in (("mixed", sg.PREC_MIXED), ("fp64", sg.PREC_FP64)):
        t = time.perf_counter()
        S = _session_qp(lib, model, arrays, q, precision=prec, use_graph=use_graph,
                        time_limit_seconds=args.solve_time_limit, autotune=0 if args.no_tune else 1)
        setup_s = time.perf_counter() - t
        S.iterate(args.warmup)
        t = time.perf_counter()
        S.iterate(args.steps)
        wall = time.perf_counter() - t
        ms = S.iterate(min(args.steps, 200), timed=True)
        r = dict(iters_per_s=args.steps / wall, kernel_ms=ms.tolist())
        if not args.skip_solve:
            loops, solves = [], []
            for _ in range(repeats):
                t = time.perf_counter()
                rc, x, y, st = S.solve()
                solves.append(time.perf_counter() - t)
                loops.append(st.loop_seconds)
            kkt = _kkt(lib, model, arrays, q, x, y)
            r.update(tiling=dict(tuned=st.tuned, tile_nnz=st.tile_nnz, spmv_grid=st.spmv_grid,
                                 probe_grid=st.probe_grid, xcache_bytes=st.xcache_bytes,
                                 stream_loads=st.stream_loads, tune_s=round(st.tune_seconds, 3)))
            r.update(converged=bool(st.converged), iterations=st.iterations, fp32_iterations=st.fp32_iterations,
                     fp64_iterations=st.fp64_iterations, loop_s=float(np.median(loops)), loop_s_all=loops,
                     end_to_end_s=setup_s + float(np.median(solves)), setup_s=setup_s, restarts=st.restarts,
                     original_kkt=kkt[:3].tolist(), objective=float(kkt[3]))
            log(f"{name}: {r['iters_per_s']:.0f} it/s, solve conv={st.converged} it={st.iterations} "
                f"loop={st.loop_seconds:.3f}s kkt={kkt[:3]}")
        S.close()
        res[name] = r
    return res


def _line(metric_value, unit, args, workload, extra):
    line = {"metric": METRIC, "value": metric_value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32 iterates anchored in fp64, fp64 checks", "data": "synthetic (seeded generator, in-process)",
            "config": {"workload": workload}}
    line.update(extra)
    return line


def workload_lp0(args):
    """The reference's CPU-runnable LP: m=2000, n=5000, 5 nnz per column."""
    lib = sg.load_gpu()
    model = sg.Model.generated("skw_bounded_lp", 2000, 5000, 5, 20260825)
    arrays = _arrays(model)
    res = _engines(lib, model, arrays, None, args, use_graph=0 if args.no_graph else 1)
    extra = {"engines": res}
    if not args.skip_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import sk_oracle as so
        kind = "reference" if so.load_ref() is not None else "port"
        fn = so.first_order_ref if kind == "reference" else so.first_order_orc
        t = time.perf_counter()
        rc, s = fn(model, 0)
        dt = time.perf_counter() - t
        extra["cpu_baseline"] = {"value": s.iterations / dt, "unit": "PDHG iterations/s", "cores": 1, "kind": kind,
                                 "sample": f"full reference CPU PDHG run: {s.iterations} iterations, {dt:.2f} s, "
                                           f"result {sg.RESULT_NAMES[s.result]}, objective {s.objective:.10g}"}
    mix = res["mixed"]
    if not args.skip_solve and mix.get("converged") and res["fp64"].get("converged"):
        extra["speedup_time_to_kkt_vs_fp64"] = round(res["fp64"]["loop_s"] / mix["loop_s"], 3)
    return _line(round(mix["iters_per_s"], 1), "PDHG iterations/s", args, "LP m=2000 n=5000 (config 0)", extra)


def workload_qp2(args):
    """Convex QP: n=500,000, m=250,000, Q = D + B'B."""
    lib = sg.load_gpu()
    t = time.perf_counter()
    model = sg.Model.generated("skw_sparse_qp", 250000, 500000, 6, 50000, 3, 20260827)
    arrays = _arrays(model)
    q = model.q_csc()
    log(f"QP built in {time.perf_counter() - t:.1f}s, nnz(A)={int(arrays[0][-1])} nnz(Q)={int(q[0][-1])}")
    res = _engines(lib, model, arrays, q, args, use_graph=0 if args.no_graph else 1)
    extra = {"engines": res, "nnz_A": int(arrays[0][-1]), "nnz_Q": int(q[0][-1])}
    mix = res["mixed"]
    if not args.skip_solve and mix.get("converged") and res["fp64"].get("converged"):
        extra["speedup_time_to_kkt_vs_fp64"] = round(res["fp64"]["loop_s"] / mix["loop_s"], 3)
    return _line(round(mix["iters_per_s"], 1), "PDHG iterations/s", args,
                 "convex QP n=500000 m=250000, A 6/col, Q=D+B'B (config 2)", extra)


def workload_batch3(args):
    """4096 small LPs in one persistent launch."""
    count = 4096 if not args.quick else 256
    w = sg.load_workloads()
    sizes = (C.c_int * count)()
    seed = 20260828
    w.skw_batch_sizes(count, 20, 400, seed, sizes)
    t = time.perf_counter()
    models = [sg.Model.generated("skw_bounded_lp", sizes[i], 2 * sizes[i], 5, seed * count + i) for i in range(count)]
    log(f"{count} instances built in {time.perf_counter() - t:.1f}s")
    rc, xs, ys, res, st = sg.batch_solve(models, eps_optimal=1e-6, max_iterations=200000)
    if rc != 0:
        raise RuntimeError(sg.last_error())
    its = np.array([r.iterations for r in res])
    conv = np.array([r.status == 0 for r in res])
    log(f"batch: device {st.device_seconds:.3f}s setup {st.setup_seconds:.2f}s converged {conv.sum()}/{count} "
        f"threads {st.block_threads} tune {list(st.tune_ms)}")
    extra = {"batch": {"instances": count, "converged": int(conv.sum()), "device_s": st.device_seconds,
                       "setup_s": st.setup_seconds, "block_threads": st.block_threads, "grid": st.grid,
                       "tune_ms": list(st.tune_ms),
                       "iterations": {"min": int(its.min()), "median": float(np.median(its)),
                                      "p90": float(np.percentile(its, 90)), "max": int(its.max()),
                                      "total": int(its.sum())}}}
    # plain FP64 GPU PDHG, one instance at a time, on a sample
    if not args.skip_solve:
        lib = sg.load_gpu()
        sample = list(range(0, count, max(1, count // 32)))
        t = time.perf_counter()
        for i in sample:
            S = _session_qp(lib, models[i], _arrays(models[i]), None, precision=sg.PREC_FP64, use_graph=1)
            S.solve()
            S.close()
        seq = (time.perf_counter() - t) / len(sample)
        extra["sequential_fp64_s_per_instance"] = seq
        extra["batch_speedup_vs_sequential_fp64"] = round(seq * count / st.device_seconds, 1)
    if not args.skip_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import sk_oracle as so
        kind = "reference" if so.load_ref() is not None else "port"
        fn = so.first_order_ref if kind == "reference" else so.first_order_orc
        t = time.perf_counter()
        done = 0
        while done < count and time.perf_counter() - t < 20.0:
            fn(models[done], 0)
            done += 1
        dt = time.perf_counter() - t
        extra["cpu_baseline"] = {"value": done / dt, "unit": "instances/s", "cores": 1, "kind": kind,
                                 "sample": f"first {done} instances through the reference CPU PDHG to its own stop"}
    return _line(round(count / st.device_seconds, 1), "LP instances/s (to 1e-6 KKT)", args,
                 "4096 bounded LPs m~U{20..400}, n=2m, 5 nnz/col (config 3)", extra)


def workload_warm4(args):
    """Learned warm start on the parametric LP family (m=20,000, n=50,000)."""
    lib = sg.load_gpu()
    w = sg.load_workloads()
    if args.train is None:
        args.train = 1000
    if args.epochs is None:
        args.epochs = 300
    m, n, dim = (20000, 50000, 16) if not args.quick else (2000, 5000, 16)
    seed = 20260829
    model = sg.Model()
    P = np.zeros(n * dim); R = np.zeros(m * dim); c0 = np.zeros(n); b0 = np.zeros(m)
    if w.skw_parametric_family(C.byref(model.m), m, n, 6, dim, seed, sg.dptr(P), sg.dptr(R), sg.dptr(c0),
                               sg.dptr(b0)) != 0:
        raise RuntimeError("family generation failed")
    model._generated = True
    P = P.reshape(n, dim); R = R.reshape(m, dim)
    th_train = np.zeros(args.train * dim); th_test = np.zeros(args.test * dim)
    w.skw_normal_matrix(args.train, dim, seed + 1, sg.dptr(th_train))
    w.skw_normal_matrix(args.test, dim, seed + 2, sg.dptr(th_test))
    th_train = th_train.reshape(args.train, dim); th_test = th_test.reshape(args.test, dim)
    l, u = model.vec("clow"), model.vec("cupp")

    def instance(theta):
        half = 0.5 + 0.1 * np.abs(R @ theta)
        return c0 + P @ theta, b0 - half, b0 + half

    arrays = _arrays(model)
    eps_train = args.train_eps
    S = _session_qp(lib, model, arrays, None, precision=sg.PREC_MIXED, use_graph=1, eps_optimal=eps_train)
    Y = np.zeros((args.train, n + m), np.float32)
    t0 = time.perf_counter()
    gen_iters = []
    xp = yp = None
    for k in range(args.train):
        c, rl, ru = instance(th_train[k])
        S.update(c, rl, ru, l, u)
        rc, x, y, st = S.solve(xp, yp)   # plain PDHG, warm-started from the previous training solution
        Y[k, :n] = x; Y[k, n:] = y
        xp, yp = x, y
        gen_iters.append(st.iterations)
        if (k + 1) % 500 == 0:
            log(f"training data {k + 1}/{args.train}: {time.perf_counter() - t0:.1f}s, "
                f"median iterations {np.median(gen_iters):.0f}")
    gen_s = time.perf_counter() - t0
    S.close()
    X = th_train.astype(np.float32)
    net = sg.Mlp(dim, 512, n + m, seed=seed)
    ysd = Y.std(0)
    net.set_scaling(X.mean(0), X.std(0), Y.mean(0), np.where(ysd > 1e-6, ysd, 1.0))
    t = time.perf_counter()
    hist = net.train(X, Y, epochs=args.epochs, batch=256, lr=1e-3)
    train_s = time.perf_counter() - t
    log(f"MLP trained in {train_s:.1f}s, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    t = time.perf_counter()
    Yhat = net.predict(th_test.astype(np.float32))
    infer_s = time.perf_counter() - t
    # cold versus warm on held-out instances, both to 1e-6 relative KKT
    S = _session_qp(lib, model, arrays, None, precision=sg.PREC_MIXED, use_graph=1, eps_optimal=1e-6,
                    time_limit_seconds=60.0)
    # the learned start gets twice the training solves' median iteration
    # count before its cold-start fallback takes over: a budget known before
    # the held-out instance is seen
    warm_budget = int(2 * np.median(gen_iters))
    cold_it, warm_it, cold_s, warm_s, errs, fallbacks, conv_c, conv_w = [], [], [], [], [], 0, 0, 0
    for k in range(min(args.compare, args.test)):
        c, rl, ru = instance(th_test[k])
        S.update(c, rl, ru, l, u)
        rc, xc, yc, stc = S.solve()
        x0 = np.clip(Yhat[k, :n].astype(np.float64), l, u)
        y0 = Yhat[k, n:].astype(np.float64)
        xw, yw, stw = np.zeros(n), np.zeros(m), sg.PdlpStats()
        used = C.c_int(0)
        if lib.sk_gpu_pdlp_session_solve_fallback(S.h, sg.dptr(x0), sg.dptr(y0), warm_budget,
                                                  sg.dptr(xw), sg.dptr(yw), C.byref(stw), C.byref(used)) != 0:
            raise RuntimeError(sg.last_error())
        if not used.value:
            fallbacks += 1
        cold_it.append(stc.iterations); warm_it.append(stw.iterations)
        # both timed the same way: start (upload of the start point) + loop
        cold_s.append(stc.seconds); warm_s.append(stw.seconds)
        conv_c += stc.converged; conv_w += stw.converged
        sol = np.concatenate([xc, yc])
        errs.append(float(np.linalg.norm(Yhat[k] - sol) / max(np.linalg.norm(sol), 1e-12)))
    S.close()
    extra = {"family": {"m": m, "n": n, "theta_dim": dim, "train": args.train, "test": args.test,
                        "compared": len(cold_it)},
             "training_data": {"seconds": gen_s, "eps": eps_train, "median_iterations": float(np.median(gen_iters))},
             "mlp": {"params": net.n_params, "epochs": args.epochs, "train_s": train_s,
                     "loss_first": float(hist[0]), "loss_last": float(hist[-1]), "inference_s_all_test": infer_s,
                     "loss_curve_every_10_epochs": [round(float(v), 4) for v in hist[::10]],
                     "median_rel_error_vs_cold_solution": float(np.median(errs))},
             "cold": {"converged": conv_c, "median_iterations": float(np.median(cold_it)),
                      "mean_seconds": float(np.mean(cold_s))},
             "warm": {"converged": conv_w, "median_iterations": float(np.median(warm_it)),
                      "mean_seconds": float(np.mean(warm_s)), "fallbacks": fallbacks,
                      "warm_iteration_budget": warm_budget},
             "speedup_time_warm_vs_cold": float(np.mean(cold_s) / max(np.mean(warm_s), 1e-12)),
             "speedup_iterations_warm_vs_cold": float(np.median(cold_it) / max(np.median(warm_it), 1))}
    log(json.dumps(extra))
    return _line(round(extra["speedup_time_warm_vs_cold"], 3), "x wall-time speedup (warm vs cold, to 1e-6)", args,
                 "parametric LP m=20000 n=50000, theta in R^16, MLP 16-512-512-70000 (config 4)", extra)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1280)
    ap.add_argument("--warmup", type=int, default=128)
    ap.add_argument("--workload", default="lp1", choices=sorted(WORKLOADS),
                    help="lp1 (default, headline) | lp0 | qp2 | batch3 | warm4 | mlp4 | pipe1")
    ap.add_argument("--skip-solve", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--quick", action="store_true", help="small instance (smoke)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch restart windows directly instead of replaying hipGraphs (profilers)")
    ap.add_argument("--solve-time-limit", type=float, default=45.0)
    ap.add_argument("--engine", default="mixed", choices=["mixed", "fp64"], help="lp1ndg: which engine")
    ap.add_argument("--skip-ndg", action="store_true", help="lp1: no NDG-restart solve")
    ap.add_argument("--no-tune", action="store_true", help="lp0, lp1, qp2, pipe1: fixed tiling instead of the autotuner")
    ap.add_argument("--tiling", default="",
                    help="lp1: fix the tiling as TILE,GRID,STREAM[,PROBE_GRID,XCACHE_BYTES] (no autotuner)")
    ap.add_argument("--train", type=int, default=None,
                    help="warm4: training instances (default 1000); mlp4: rows (default 4096)")
    ap.add_argument("--test", type=int, default=100, help="warm4: held-out instances")
    ap.add_argument("--compare", type=int, default=20, help="warm4: held-out instances solved cold and warm")
    ap.add_argument("--epochs", type=int, default=None, help="warm4: default 300; mlp4: default 2")
    ap.add_argument("--train-eps", type=float, default=1e-6, help="warm4: tolerance of the training solves")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="copy this many MiB device-to-device 3 times first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        hip = sg._hip_runtime()
        a, b = C.c_void_p(), C.c_void_p()
        hip.hipMalloc(C.byref(a), nbytes)
        hip.hipMalloc(C.byref(b), nbytes)
        for _ in range(3):
            hip.hipMemcpy(b, a, nbytes, 3)
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
        log(f"copy calibration: 3 x {nbytes} bytes device-to-device")
    line = WORKLOADS[args.workload](args)
    print(json.dumps(line), flush=True)


def workload_mlp4(args):
    """The warm-start MLP's kernels alone at the config-4 shape
    (16 -> 512 -> 512 -> 70,000, batch 256) on random standardised data:
    training and inference throughput against the BF16 MFMA peak."""
    n_in, hid, out = 16, 512, 70000
    rows = max(args.train if args.train is not None else 4096, 512)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((rows, n_in)).astype(np.float32)
    Y = rng.standard_normal((rows, out)).astype(np.float32)
    net = sg.Mlp(n_in, hid, out, seed=1)
    net.set_scaling(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    net.train(X[:512], Y[:512], epochs=1, batch=256, lr=1e-3)      # warm-up
    epochs = max(args.epochs if args.epochs is not None else 2, 1)
    # a training call uploads the whole set first (rows x 70,000 floats):
    # timed alone with no epochs and taken off, so the rate is the steps'
    t = time.perf_counter()
    net.train(X, Y, epochs=0, batch=256, lr=1e-3)
    upload_s = time.perf_counter() - t
    t = time.perf_counter()
    net.train(X, Y, epochs=epochs, batch=256, lr=1e-3)
    call_s = time.perf_counter() - t
    train_s = max(call_s - upload_s, 1e-9)
    steps = epochs * (rows // 256)
    per_sample = 2 * (n_in * hid + hid * hid + hid * out)             # forward FLOP
    train_tf = steps * 256 * 3 * per_sample / train_s / 1e12
    t = time.perf_counter()
    net.predict(X)
    infer_s = time.perf_counter() - t
    extra = {"mlp": {"params": net.n_params, "rows": rows, "epochs": epochs, "steps": steps,
                     "ms_per_step": round(1e3 * train_s / steps, 3), "train_TFLOP_s": round(train_tf, 1),
                     "upload_and_setup_s": round(upload_s, 4), "call_s_incl_upload": round(call_s, 4),
                     "predict_s": round(infer_s, 4),
                     "predict_TFLOP_s_incl_transfers": round(rows * per_sample / infer_s / 1e12, 2)},
             "roofline": {"bound": "mfma", "achieved": round(train_tf, 1), "peak": 2500.0, "unit": "TFLOP/s",
                          "frac": round(train_tf / 2500.0, 4), "traffic": None,
                          "note": "whole training step (GEMMs, transposes, Adam) over dense BF16 peak"}}
    line = _line(round(train_tf, 1), "TFLOP/s (MLP training step)", args,
                 "MLP 16-512-512-70000 BF16 MFMA, batch 256, random data (config-4 shape)", extra)
    line["dtype"] = "bf16 MFMA, fp32 accumulate"
    return line


def workload_pipe1(args):
    """The headline LP through the LP pipeline (sk_gpu_solve_lp: host presolve,
    GPU PDLP with FP32 iterates on the reduced model, postsolve, original-space
    gate), with the reference's own presolve (oracle/_ref) timed beside ours on
    the same problem on the host."""
    t = time.perf_counter()
    model, _ = make_workload(args.quick)
    problem = sg.LpProblem.from_model(model)
    log(f"LpProblem built in {time.perf_counter() - t:.1f}s ({problem.rows} x {problem.cols}, nnz {problem.nnz})")
    opts = sg.lp_solver_options(use_presolve=1, eps_optimal=1e-6, time_limit_seconds=args.solve_time_limit,
                                precision=sg.PREC_MIXED, use_graph=0 if args.no_graph else 1,
                                autotune=0 if args.no_tune else 1)
    t = time.perf_counter()
    sol, _x = sg.solve_lp(problem, opts)
    wall = time.perf_counter() - t
    st = sol.pdlp
    log(f"pipeline: {sg.LP_STATUS_NAMES[sol.status]} presolve {sol.presolve_seconds:.2f}s "
        f"solve {sol.solve_seconds:.2f}s iterations {sol.iterations}")
    extra = {"pipeline": {"status": sg.LP_STATUS_NAMES[sol.status], "converged": bool(st.converged),
                          "presolve_s": round(sol.presolve_seconds, 3), "solve_s": round(sol.solve_seconds, 3),
                          "wall_s": round(wall, 3), "iterations": sol.iterations,
                          "fp32_iterations": st.fp32_iterations, "fp64_iterations": st.fp64_iterations,
                          "removed_rows": sol.presolve_removed_rows, "removed_cols": sol.presolve_removed_cols,
                          "kkt_reduced": [st.relative_primal_residual, st.relative_dual_residual, st.relative_gap],
                          "original_space_primal_residual": sol.primal_residual,
                          "time_limit_s": args.solve_time_limit}}
    if not args.skip_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import sk_oracle as so
        if so.load_presolve_ref() is not None:
            t = time.perf_counter()
            ref = so.ref_presolve(problem)
            dt = time.perf_counter() - t
            extra["cpu_baseline"] = {"value": round(dt, 3), "unit": "s (presolve)", "cores": 1, "kind": "reference",
                                     "sample": "the reference's presolve (src/lp/Presolve.cpp, compiled from its "
                                               f"sources) on the same problem, whole run; removed "
                                               f"{ref['removed_rows']} rows and {ref['removed_cols']} columns"}
            extra["presolve_speedup_vs_reference"] = round(dt / max(sol.presolve_seconds, 1e-9), 3)
    rate = sol.iterations / max(st.loop_seconds, 1e-9) if sol.iterations else 0.0
    line = _line(round(rate, 1), "PDHG iterations/s", args,
                 "LP m=1000000 n=2000000 through presolve + PDLP + postsolve (config 1 pipeline)", extra)
    line["higher_is_better"] = True
    return line


def workload_lp1ndg(args):
    """The headline LP to relative KKT 1e-6 with NDG restarts -- the one
    restart criterion under which it gets there -- for one engine
    (--engine mixed|fp64): one solve, the session resumed every 60 s (each
    resume prints a progress line), up to --solve-time-limit seconds.
    Minutes long: not part of the default run."""
    lib = sg.load_gpu()
    model, _ = make_workload(args.quick)
    rp, cj, v, vec = _arrays(model)
    prec = sg.PREC_FP64 if args.engine == "fp64" else sg.PREC_MIXED
    S = Session(lib, model, (rp, cj, v, vec), precision=prec, restart_rule=sg.RESTART_NDG,
                use_graph=0 if args.no_graph else 1, autotune=0 if args.no_tune else 1)
    chunk = min(60.0, args.solve_time_limit)
    S.set_limits(100_000_000, chunk, 1e-6)
    t0 = time.perf_counter()
    rc, x, y, st = S.solve()
    while rc == 0 and not st.converged and st.loop_seconds < args.solve_time_limit:
        log(f"{args.engine} NDG: {st.iterations} iterations, {st.loop_seconds:.0f} s, relative KKT "
            f"{st.relative_primal_residual:.2e} {st.relative_dual_residual:.2e} {st.relative_gap:.2e}")
        x = np.zeros(model.ncol); y = np.zeros(model.nrow); st = sg.PdlpStats()
        rc = lib.sk_gpu_pdlp_session_resume(S.h, sg.dptr(x), sg.dptr(y), C.byref(st))
    wall = time.perf_counter() - t0
    kkt = np.zeros(5)
    lib.sk_gpu_relative_kkt(model.nrow, model.ncol, int(rp[-1]), sg.iptr(rp), sg.iptr(cj), sg.dptr(v),
                            sg.dptr(vec["c"]), sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]), sg.dptr(vec["clow"]),
                            sg.dptr(vec["cupp"]), sg.dptr(x), sg.dptr(y), sg.dptr(kkt))
    d = st.as_dict()
    S.close()
    return _line(round(st.iterations / max(st.loop_seconds, 1e-9), 1), "PDHG iterations/s", args,
                 f"LP {model.nrow} x {model.ncol} (power-law generator) to 1e-6, NDG restarts",
                 {"engine": args.engine, "converged": bool(st.converged), "iterations": st.iterations,
                  "time_to_kkt_1e-6_s": st.loop_seconds if st.converged else None, "wall_s": wall,
                  "seconds_to_merit": d["seconds_to_merit"], "restarts": st.restarts,
                  "original_kkt": kkt[:3].tolist()})


WORKLOADS = {"lp1": workload_lp1, "lp0": workload_lp0, "qp2": workload_qp2, "batch3": workload_batch3,
             "warm4": workload_warm4, "mlp4": workload_mlp4, "pipe1": workload_pipe1,
             "lp1ndg": workload_lp1ndg}

if __name__ == "__main__":
    main()
