"""Presolve fixtures written by the reference's own presolve (src/lp/
Presolve.cpp, compiled from the reference sources with oracle/ref_presolve.cpp
into oracle/_ref by `make -C oracle ref`).

    python tests/golden/make_presolve_golden.py [reference data dir]

Writes tests/golden/presolve/<model>.npz for every MPS model of the
reference's data/miplib2017_small (read by the reference's own reader,
read_mps_file + lp_problem_from_mps, integrality ignored), and
tests/golden/presolve/random.npz with 200 seeded random LPs
(presolve_cases.random_lp, seed 2024), packed by presolve_cases.pack.  Each
case holds the problem under in/<field> and the reference's result under
out/<field>."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
for p in ("the-solver_amd", "oracle", "tests"):
    sys.path.insert(0, str(ROOT / p))
import presolve_cases as pc  # noqa: E402
import sk_oracle as so  # noqa: E402

OUT = ROOT / "tests" / "golden" / "presolve"
RANDOM_CASES, RANDOM_SEED = 200, 2024


def case(problem, prefix=""):
    d = {f"{prefix}in/{k}": np.asarray(v) for k, v in problem.arrays().items()}
    res = so.ref_presolve(problem)
    d.update({f"{prefix}out/{k}": v for k, v in pc.result_arrays(res).items()})
    return d, res


def main():
    if so.load_presolve_ref() is None:
        raise SystemExit("oracle/_ref is not built: make -C oracle ref")
    data = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference/data/miplib2017_small")
    OUT.mkdir(parents=True, exist_ok=True)
    for path in sorted(data.glob("*.mps")):
        d, res = case(so.ref_lp_from_mps(path))
        np.savez_compressed(OUT / f"{path.stem}.npz", **d)
        print(path.stem, "status", res["status"], "removed", res["removed_rows"], res["removed_cols"])
    rng = np.random.default_rng(RANDOM_SEED)
    cases = [case(pc.random_lp(rng))[0] for _ in range(RANDOM_CASES)]
    np.savez_compressed(OUT / "random.npz", **pc.pack(cases))
    print("random", RANDOM_CASES)


if __name__ == "__main__":
    main()
