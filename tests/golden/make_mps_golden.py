"""Expected models for the MPS / QPS fixtures under tests/golden/mps/,
written by the reference C core's own reader (sk_read_mps, compiled from
the reference sources into oracle/_ref by `make -C oracle ref`).

    python tests/golden/make_mps_golden.py

Writes tests/golden/mps_expected.npz: for each fixture file F, the keys
F/status and, when it parsed, F/<field> for every sk_model field."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "the-solver_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
import sk_oracle as so  # noqa: E402


def main():
    if so.load_ref() is None:
        raise SystemExit("oracle/_ref is not built: make -C oracle ref")
    out = {}
    for path in sorted((ROOT / "tests" / "golden" / "mps").iterdir()):
        st, arrays = so.ref_read_mps(path)
        out[f"{path.name}/status"] = np.array(st)
        if arrays is not None:
            for k, v in arrays.items():
                out[f"{path.name}/{k}"] = np.asarray(v)
        print(path.name, "status", st)
    np.savez_compressed(ROOT / "tests" / "golden" / "mps_expected.npz", **out)


if __name__ == "__main__":
    main()
