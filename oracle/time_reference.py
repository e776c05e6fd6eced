"""Time the reference program against the C oracle on this container's CPU.

Runs in the development container only: oracle/_ref/sam (`make -C oracle
ref-sam`) under the image's MPICH mpirun with `--np` ranks, and the oracle
with the same number of OpenMP threads, on the stratified DNS workload (seeded
random IC, N = 4, nu = 5e-4, fixed dt) at each grid of `--n`.  The per-step
time is (t(n_steps = a + b) - t(n_steps = a)) / b, which removes start-up
and I/O.  Prints one JSON line per grid.

    python oracle/time_reference.py --np 8 --n 64 128 256
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path = [ROOT] + [q for q in sys.path if os.path.abspath(q or ".") != HERE]

from oracle import oracle as O  # noqa: E402
from sam_amd import ic  # noqa: E402
from sam_amd import io as sio  # noqa: E402

SAM = os.path.join(HERE, "_ref", "sam")
MPIRUN = "/opt/conda/bin/mpirun"


def ref_time(n, np_, steps, u0, params):
    d = os.path.join(HERE, "_ref", "runs", f"time_{n}_{steps}")
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(d)
    sio.write_vel(os.path.join(d, "vel.000000"), u0)
    sio.write_header(os.path.join(d, "header.000000"), params, time=0.0, dt=params["dt0"], nt=0)
    sio.write_input(os.path.join(d, "input.dat"), **dict(params, n_steps=steps))
    t0 = time.perf_counter()
    r = subprocess.run([MPIRUN, "-np", str(np_), SAM], cwd=d, capture_output=True, text=True, timeout=3600)
    dt = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    shutil.rmtree(d)
    return dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--np", type=int, default=8)
    ap.add_argument("--n", type=int, nargs="+", default=[64, 128, 256])
    ap.add_argument("--a", type=int, default=1)
    ap.add_argument("--b", type=int, default=4)
    args = ap.parse_args()
    for n in args.n:
        params = dict(Nx=n, Ny=n, Nz=n, xL=2 * np.pi, yL=2 * np.pi, zL=2 * np.pi, vis=5e-4, i_restart=1,
                      nt_restart=0, i_strat=1, grav=16.0, lapse0=1.0, To=1.0, Pr=1.0, z0=0.0, i_cfl=0,
                      dt0=0.002, nrk_max=3, n_dealias=1, i_fs=0)
        u0 = ic.random_solenoidal(n, k0=4.0, seed=20261203, i_strat=1)
        ta = ref_time(n, args.np, args.a, u0, params)
        tb = ref_time(n, args.np, args.a + args.b, u0, params)
        t_ref = (tb - ta) / args.b
        o = O.Oracle(O.make_cfg(n, i_strat=1, grav=16.0, lapse0=1.0, vis=5e-4, i_cfl=0, dt0=0.002), threads=args.np)
        o.step(u0, 1)
        t0 = time.perf_counter()
        o.step(u0, args.b)
        t_orc = (time.perf_counter() - t0) / args.b
        print(json.dumps({"n": n, "ranks_threads": args.np, "reference_s_per_step": round(t_ref, 4),
                          "oracle_s_per_step": round(t_orc, 4), "oracle_over_reference": round(t_orc / t_ref, 3),
                          "reference_Mpt_steps_s": round(n ** 3 / t_ref / 1e6, 3)}), flush=True)


if __name__ == "__main__":
    main()
