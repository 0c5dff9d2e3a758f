"""Writes tests/golden/bkw_n16_steps.npz: the homogeneous BKW relaxation of
Maxwell molecules stepped by the FP64 oracle.  16^3 on [-6, 6]^3, M = 6,
Nrho = 8, kernel 1/(4 pi), from the BKW state at t0 = 5.5:
  f_rk2   -- after 100 SSP-RK2 (Heun) steps of dt = 0.01 (t = 6.5),
  f_euler -- after 10 forward-Euler steps of dt = 0.01 (t = 5.6).
The stage arithmetic is that of frfs/integrators/std/steppers.py:39-225 with
no advection.  Run from the repo root:
    python tests/golden/make_bkw_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import vhs as O  # noqa: E402

N, L, NRHO, M, T0, DT = 16, 6.0, 8, 6, 5.5, 0.01


def oracle_run(scheme, nsteps):
    vm = O.VelocityMesh(N, L, NRHO, M)
    orc = O.VHSOracle(vm, 0.0, 1.0 / (4.0 * math.pi))
    f = O.bkw(vm, T0)
    for _ in range(nsteps):
        if scheme == "euler":
            f = f + DT * orc.q(f)
        else:
            f1 = f + DT * orc.q(f)
            f = 0.5 * f + 0.5 * f1 + 0.5 * DT * orc.q(f1)
    return f


def main():
    out = dict(f_rk2=oracle_run("rk2", 100), f_euler=oracle_run("euler", 10),
               dims=np.array([N, NRHO, M]), params=np.array([L, T0, DT]))
    np.savez_compressed(os.path.join(HERE, "bkw_n16_steps.npz"), **out)
    print("bkw_n16_steps.npz written")


if __name__ == "__main__":
    main()
