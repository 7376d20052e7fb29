"""Writes ghost_stream.json: entries of the ghost stream from the oracle
(oracle/ghost.py, over its C restatement), for seed 0xdeadbeef and seed 1,
rows 0..63 and 1,000,000..1,000,015 of columns 1, 2 and 300, in ComplexF64
and Float64. The fixture pins this backend's own stream (column values that
depend on (seed, col) alone) against later changes, on the host and on the
device. Run from the repository root: python tests/golden/make_ghost_fixture.py"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import ghost as OG  # noqa: E402

CASES = [(0xDEADBEEF, 0, 64), (0xDEADBEEF, 1000000, 16), (1, 0, 64)]
COLS = (1, 2, 300)


def main():
    out = {"cols": list(COLS), "cases": []}
    for seed, row0, n in CASES:
        case = {"seed": seed, "row0": row0, "rows": n, "complex128": {}, "float64": {}}
        for col in COLS:
            z = OG.ghost(np.complex128, row0, n, col, 1, seed)[:, 0]
            x = OG.ghost(np.float64, row0, n, col, 1, seed)[:, 0]
            case["complex128"][str(col)] = [[float(v.real), float(v.imag)] for v in z]
            case["float64"][str(col)] = [float(v) for v in x]
        out["cases"].append(case)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ghost_stream.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(path)


if __name__ == "__main__":
    main()
