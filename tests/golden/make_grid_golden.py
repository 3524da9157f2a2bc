"""Writes tests/golden/reference_grid.json: the metrics and Coriolis
parameter the reference's own grid set-up computes (Service/
basin_parameters.f90, basinpar, for the shipped Inc/basinpar.fi), through
the compiled reference harness's mode 8 (oracle/_ref, built by
make -C oracle ref where the reference sources are mounted).  Each field is
recorded as its SHA-256 (little-endian float64 bytes, (ny, nx) C order),
L2 norm and a few values.

    python tests/golden/make_grid_golden.py
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "inmsom-elfs_amd"), os.path.join(ROOT, "oracle"), HERE]

import oracle  # noqa: E402
from inmsom_bt.cases import REF_NX, REF_NY, make_basin, write_case  # noqa: E402
from make_golden import digest  # noqa: E402

OUT = os.path.join(HERE, "reference_grid.json")
PROBES = [(2, 2), (80, 100), (160, 286), (1, 1), (0, 0)]  # (row, column), 0-based


def main():
    if not oracle.have_reference():
        sys.exit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    b = make_basin(REF_NX - 4, REF_NY - 4, seed=1, sigma=4.0)
    with tempfile.TemporaryDirectory() as tmp:
        case = os.path.join(tmp, "case.bin")
        write_case(case, b, np.zeros((b.ny, b.nx)), 1, 1, 0, 8, [])
        ref = oracle.run_reference(case, case + ".out", b.nx, b.ny, 8)
    out = {"generator": "oracle/_ref/ref_harness mode 8 (basinpar of the shipped Inc/basinpar.fi)",
           "nx": b.nx, "ny": b.ny, "probes": PROBES,
           "fields": {k: dict(digest(v), probe_values=[float(v[r, c]) for r, c in PROBES])
                      for k, v in ref.items()}}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
