#!/usr/bin/env python3
"""Benchmark: restarted PDHG (PDLP) on one MI355X.

Workload (the headline configuration): bounded LP with m = 1,000,000 rows and
n = 2,000,000 columns, Zipf(1.2) column degrees with mean 8 capped at 2000,
log-uniform |a_ij| in [1e-3, 1e3], 30% equality rows, bounds [0, 10]
(generator skw_zipf_lp, seed 20260826), synthetic and generated in-process.

A "step" is one PDHG iteration of the adaptive-step loop (dual prox, A^T y
with fused primal prox, A x with fused step decision) on the resident
problem.  `value` is PDHG iterations per second with FP32 iterates (the
mixed-precision engine); the FP64 engine ("plain FP64 GPU PDHG") is timed
the same way beside it.  Unless --skip-solve, both engines also solve the LP
from a cold start to a relative KKT of 1e-6 and the wall times are reported.
The roofline object reports the dominant kernel's algorithmic HBM bytes
divided by its HIP-event-timed mean duration on the library stream.

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
        ms = np.zeros(4)
        rc = self.lib.sk_gpu_pdlp_session_iterate(self.h, k, sg.dptr(ms) if timed else None)
        if rc != 0:
            raise RuntimeError(f"iterate failed: {sg.last_error()}")
        return ms

    def solve(self):
        x = np.zeros(self.model.ncol)
        y = np.zeros(self.model.nrow)
        st = sg.PdlpStats()
        rc = self.lib.sk_gpu_pdlp_session_solve(self.h, None, None, sg.dptr(x), sg.dptr(y), C.byref(st))
        return rc, x, y, st

    def close(self):
        if self.h:
            self.lib.sk_gpu_pdlp_session_destroy(self.h)
            self.h = C.c_void_p()


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
    sk_solve(lp_engine = FIRST_ORDER) for a bounded iteration budget."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import sk_oracle as so

    kind = "reference" if so.load_ref() is not None else "port"
    fn = so.first_order_ref if kind == "reference" else so.first_order_orc
    t0 = time.perf_counter()
    rc, s = fn(model, iters)
    dt = time.perf_counter() - t0
    done = int(s.iterations)
    return {"value": done / dt, "unit": "PDHG iterations/s", "cores": 1, "kind": kind,
            "sample": f"{done} iterations of the reference CPU PDHG (sk_solve, lp_engine=FIRST_ORDER, "
                      f"threads=1) on the same LP, {dt:.1f} s incl. its setup"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--skip-solve", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=100)
    ap.add_argument("--quick", action="store_true", help="small instance (smoke)")
    ap.add_argument("--solve-time-limit", type=float, default=150.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    lib = sg.load_gpu()
    t0 = time.perf_counter()
    model, q = make_workload(args.quick)
    rp, cj, v, vec = None, None, None, None
    p, i, x = model.csc()
    rp, cj, v = sg.csc_to_csr(model.nrow, model.ncol, p, i, x)
    vec = {k: model.vec(k) for k in ("c", "rlow", "rupp", "clow", "cupp")}
    arrays = (rp, cj, v, vec)
    m, n, nnz = model.nrow, model.ncol, int(rp[-1])
    log(f"workload m={m} n={n} nnz={nnz} zipf_q={q:.6f} built in {time.perf_counter() - t0:.1f}s")

    results = {}
    for name, prec in (("mixed", sg.PREC_MIXED), ("fp64", sg.PREC_FP64)):
        t = time.perf_counter()
        S = Session(lib, model, arrays, precision=prec, time_limit_seconds=args.solve_time_limit)
        setup = time.perf_counter() - t
        S.iterate(args.warmup)
        t = time.perf_counter()
        S.iterate(args.steps)            # cold start + K iterations, synchronised at the end
        wall = time.perf_counter() - t
        ms = S.iterate(args.steps, timed=True)
        r = dict(setup_s=setup, iters_per_s=args.steps / wall, ms_per_iter=1e3 * wall / args.steps,
                 kernel_ms=ms.tolist())
        log(f"{name}: {r['iters_per_s']:.1f} it/s, kernels(ms) {np.round(ms, 4).tolist()}")
        if not args.skip_solve:
            rc, xs, ys, st = S.solve()
            kkt = np.zeros(5)
            lib.sk_gpu_relative_kkt(m, n, nnz, sg.iptr(rp), sg.iptr(cj), sg.dptr(v), sg.dptr(vec["c"]),
                                    sg.dptr(vec["rlow"]), sg.dptr(vec["rupp"]), sg.dptr(vec["clow"]),
                                    sg.dptr(vec["cupp"]), sg.dptr(xs), sg.dptr(ys), sg.dptr(kkt))
            r.update(solve_rc=rc, converged=bool(st.converged), iterations=st.iterations,
                     fp32_iterations=st.fp32_iterations, fp64_iterations=st.fp64_iterations,
                     loop_s=st.loop_seconds, solve_s=st.seconds, restarts=st.restarts,
                     host_syncs=st.host_syncs, scaled_kkt=max(st.relative_primal_residual,
                                                              st.relative_dual_residual, st.relative_gap),
                     original_kkt=kkt[:3].tolist(), objective=float(vec["c"] @ xs))
            log(f"{name} solve: rc={rc} conv={st.converged} it={st.iterations} loop={st.loop_seconds:.2f}s "
                f"kkt={kkt[:3]}")
        S.close()
        results[name] = r

    mix, f64 = results["mixed"], results["fp64"]
    k1, k2, k3 = kernel_bytes(m, n, nnz, 4, 4, 1.0)
    kernels = {"dual_prox": (k1, mix["kernel_ms"][0]), "AT_spmv_primal": (k2, mix["kernel_ms"][1]),
               "A_spmv_probe": (k3, mix["kernel_ms"][2])}
    dom = max(kernels, key=lambda k: kernels[k][1])
    bytes_dom, ms_dom = kernels[dom]
    achieved = bytes_dom / (ms_dom * 1e-3) / 1e9
    spmv_gbs = {k: kernels[k][0] / (kernels[k][1] * 1e-3) / 1e9 for k in kernels}
    line = {
        "metric": METRIC,
        "value": round(mix["iters_per_s"], 2),
        "unit": "PDHG iterations/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mix["ms_per_iter"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 iterates + fp64 checks/tail",
        "data": "synthetic (seeded generator, in-process)",
        "config": {"workload": f"bounded LP m={m} n={n} nnz={nnz}, Zipf(1.2) column degrees mean 8 cap 2000, "
                               f"|a| log-uniform [1e-3,1e3], 30% equality rows, bounds [0,10]",
                   "engine": "adaptive-step restarted PDHG, FP32 iterates"},
        "fp64_iters_per_s": round(f64["iters_per_s"], 2),
        "fp64_ms_per_step": round(f64["ms_per_iter"], 4),
        "speedup_iters_vs_fp64": round(mix["iters_per_s"] / f64["iters_per_s"], 3),
        "spmv_gbs": {k: round(val, 1) for k, val in spmv_gbs.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes": int(bytes_dom), "mean_ms": round(ms_dom, 5)},
    }
    if not args.skip_solve:
        line["time_to_kkt_1e-6_s"] = {"mixed": mix.get("loop_s"), "fp64": f64.get("loop_s")}
        line["solve"] = {"mixed": {k: mix[k] for k in ("converged", "iterations", "fp32_iterations",
                                                        "fp64_iterations", "loop_s", "original_kkt",
                                                        "objective", "restarts")},
                         "fp64": {k: f64[k] for k in ("converged", "iterations", "loop_s", "original_kkt",
                                                      "objective", "restarts")}}
        if mix.get("converged") and f64.get("converged"):
            line["speedup_time_to_kkt_vs_fp64"] = round(f64["loop_s"] / mix["loop_s"], 3)
    if not args.skip_cpu:
        log("CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(model, args.cpu_iters)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
