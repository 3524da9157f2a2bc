"""CPU rates of the reference's own Fortran hot path and of the C
restatement on the same basin.

TEST INFRASTRUCTURE ONLY: a measurement recipe beside the oracle (bench.py's
cpu_baseline leg calls it).  The reference is compiled for its shipped grid
(Inc/basinpar.fi: 289 x 163), so both run the shipped Black Sea basin of
that size (tests/golden/blacksea: mask, bathymetry, longitude-latitude
metrics) from a random state.  The Fortran runs through oracle/_ref/
ref_harness mode 9: each outer step is sea_surface_fluxes,
sea_bottom_fluxes, the RHS assembly of Control/dyncall.f90:114-126 and
barotropic_dynamics (Function/vel_ssh.f90:379-679), timed with the
reference's own start_timer / end_timer (mpi_wtime) around the loop, one
MPI rank (the harness reads whole arrays; the compiled grid is not cut into
rank blocks) and OMP_NUM_THREADS threads.  The C restatement runs the same
steps (bto_model_step) in a child process with the same thread count.
Both are credited 2 nstep leap-frog substeps per outer step (nstep
state-advancing + nstep look-ahead), as bench.py's cpu_baseline is.

    python3 oracle/time_reference.py [--steps N] [--threads 1,2,4,8]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(ROOT, "inmsom-elfs_amd")]

import oracle  # noqa: E402
from inmsom_bt.cases import randomize_state, reference_basin, write_case  # noqa: E402

BLACKSEA = os.path.join(ROOT, "tests", "golden", "blacksea")
NSTEP = 2


def basin():
    b = reference_basin(os.path.join(BLACKSEA, "mask.txt"), BLACKSEA)
    return randomize_state(b, 3, amp_ssh=0.2, amp_uv=0.1)


def fortran_seconds(b, rdis, steps: int, threads: int) -> float:
    """Seconds of `steps` outer steps by the reference's own timer."""
    with tempfile.TemporaryDirectory() as d:
        case, out = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
        write_case(case, b, rdis, steps, NSTEP, 0, mode=9)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run([oracle.REF_HARNESS, case, out], check=True, env=env, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE, cwd=d)
        raw = np.fromfile(out, dtype=np.uint8)
        tail = raw[7 * b.nx * b.ny * 8:]
        nsteps, nthr = np.frombuffer(tail[:8].tobytes(), dtype="<i4")
        sec = float(np.frombuffer(tail[8:16].tobytes(), dtype="<f8")[0])
        assert nsteps == steps and nthr == threads, (nsteps, nthr)
        return sec


_C_CHILD = r"""
import os, sys, time
sys.path[:0] = [sys.argv[1], sys.argv[2]]
import oracle, time_reference as tr
b = tr.basin()
og = oracle.OracleGrid(b)
rdis = og.rayleigh()
t0 = time.perf_counter()
oracle.run(b, int(sys.argv[3]), tr.NSTEP, out_steps=[], rdis=rdis, grid=og)
print(time.perf_counter() - t0, int(oracle.lib().bto_num_threads()))
"""


def c_seconds(steps: int, threads: int) -> float:
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-c", _C_CHILD, HERE, os.path.join(ROOT, "inmsom-elfs_amd"), str(steps)],
                       check=True, env=env, capture_output=True, text=True)
    sec, nthr = r.stdout.split()
    assert int(nthr) == threads, (nthr, threads)
    return float(sec)


def measure(steps: int, threads_list) -> dict:
    b = basin()
    rdis = oracle.OracleGrid(b).rayleigh()
    subs = 2 * NSTEP
    work = b.n_wet * steps * subs
    res = {"basin": f"{b.nx}x{b.ny} Black Sea (shipped mask and bathymetry), {b.n_wet} wet cells",
           "outer_steps": steps, "nstep_barotrop": NSTEP, "substeps_per_outer_step": subs,
           "unit": "wet-cell steps/s", "ranks": 1, "threads": {}}
    for t in threads_list:
        fs = fortran_seconds(b, rdis, steps, t)
        cs = c_seconds(steps, t)
        res["threads"][str(t)] = {"fortran_s": fs, "c_s": cs, "fortran_rate": work / fs, "c_rate": work / cs,
                                  "c_over_fortran": fs / cs}
    best_f = max(v["fortran_rate"] for v in res["threads"].values())
    best_c = max(v["c_rate"] for v in res["threads"].values())
    res["best_fortran_rate"], res["best_c_rate"] = best_f, best_c
    res["c_over_fortran_best"] = best_c / best_f
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--threads", default="1,2,4,8")
    a = ap.parse_args()
    if not oracle.have_reference():
        sys.exit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    print(json.dumps(measure(a.steps, [int(x) for x in a.threads.split(",")]), indent=1))


if __name__ == "__main__":
    main()
