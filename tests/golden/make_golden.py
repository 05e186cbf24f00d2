"""Write the golden fixtures under tests/golden/ from the reference C core.

Runs in the development container only: it needs oracle/_ref/libsankhya_ref.so,
which oracle/Makefile compiles from the reference's own C sources.  The
fixtures are data (inputs identified by generator + seed + checksum, and the
reference's outputs); the reference itself never leaves this container.

  python tests/golden/make_golden.py [--skip-cfg0-simplex]

Fixtures:
  small_lp_<seed>.npz   bounded LPs (m<=200): reference simplex x, y, objective
                        and sk_verify residuals; reference CPU PDHG
                        (sk_solve, lp_engine=FIRST_ORDER) x, y, iterations,
                        result at a fixed iteration limit
  cfg0_simplex.npz      config-0 LP (2000 x 5000, seed 20260825): reference
                        simplex x, y, objective and sk_verify residuals
  cfg0_pdhg.npz         config-0 LP: reference CPU PDHG after 400 iterations
  qp_small.npz          a convex QP (n=600 > 512, nonempty A) through the
                        reference first-order QP path
"""
from __future__ import annotations

import argparse
import hashlib
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "the-solver_amd"))
import sankhya_gpu as sg  # noqa: E402
import sk_oracle as so  # noqa: E402

OUT = Path(__file__).resolve().parent
SMALL = [(20, 40, 101, 20000), (50, 100, 102, 2000), (80, 160, 103, 5000),
         (120, 240, 104, 3000), (200, 400, 105, 3000)]


def model_digest(model) -> str:
    h = hashlib.sha256()
    for a in model.csc():
        h.update(np.ascontiguousarray(a).tobytes())
    for k in ("c", "rlow", "rupp", "clow", "cupp"):
        h.update(model.vec(k).tobytes())
    q = model.q_csc()
    if q is not None:
        for a in q:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def simplex_fields(model):
    rc, s = so.simplex_ref(model)
    x, y = so.solution_arrays(s)
    return dict(simplex_rc=rc, simplex_result=s.result, simplex_objective=s.objective,
                simplex_x=x, simplex_y=y if y is not None else np.zeros(0),
                simplex_pinf=s.primal_infeasibility, simplex_dinf=s.dual_infeasibility,
                simplex_comp=s.complementarity, simplex_iterations=s.iterations)


def pdhg_fields(model, limit):
    rc, s = so.first_order_ref(model, limit)
    x, y = so.solution_arrays(s)
    return dict(pdhg_rc=rc, pdhg_result=s.result, pdhg_objective=s.objective,
                pdhg_x=x, pdhg_y=y if y is not None else np.zeros(0),
                pdhg_iterations=s.iterations, pdhg_limit=limit,
                pdhg_pinf=s.primal_infeasibility)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-cfg0-simplex", action="store_true")
    args = ap.parse_args()
    if so.load_ref() is None:
        raise SystemExit("oracle/_ref/libsankhya_ref.so is not built (make -C oracle ref)")
    for m, n, seed, limit in SMALL:
        model = sg.Model.generated("skw_bounded_lp", m, n, 5, seed)
        f = dict(generator="skw_bounded_lp", m=m, n=n, per_col=5, seed=seed,
                 digest=model_digest(model))
        f.update(simplex_fields(model))
        f.update(pdhg_fields(model, limit))
        np.savez(OUT / f"small_lp_{seed}.npz", **f)
        print("small", m, n, seed, f["simplex_objective"], f["pdhg_iterations"], flush=True)
    model = sg.Model.generated("skw_bounded_lp", 2000, 5000, 5, 20260825)
    base = dict(generator="skw_bounded_lp", m=2000, n=5000, per_col=5, seed=20260825,
                digest=model_digest(model))
    f = dict(base)
    f.update(pdhg_fields(model, 400))
    np.savez(OUT / "cfg0_pdhg.npz", **f)
    print("cfg0 pdhg", f["pdhg_iterations"], f["pdhg_objective"], flush=True)
    qp = sg.Model.generated("skw_sparse_qp", 300, 600, 3, 200, 3, 777)
    f = dict(generator="skw_sparse_qp", m=300, n=600, a_per_col=3, b_rows=200, b_per_row=3,
             seed=777, digest=model_digest(qp))
    f.update(pdhg_fields(qp, 20000))
    np.savez(OUT / "qp_small.npz", **f)
    print("qp", f["pdhg_rc"], f["pdhg_result"], f["pdhg_iterations"], f["pdhg_objective"], flush=True)
    if not args.skip_cfg0_simplex:
        t = time.time()
        f = dict(base)
        f.update(simplex_fields(model))
        f["simplex_seconds"] = time.time() - t
        np.savez(OUT / "cfg0_simplex.npz", **f)
        print("cfg0 simplex", f["simplex_objective"], f["simplex_seconds"], flush=True)


if __name__ == "__main__":
    main()
