"""Writes tests/golden/vhs_*.npz: (f, Q) pairs of the FP64 oracle at small
grids, used as fixed vectors by the GPU parity tests.  Run from the repo root:
    python tests/golden/make_vhs_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import vhs as O  # noqa: E402

CASES = [
    # name, N, L, Nrho, M, gamma, prefac, kind
    ("n4_rand", 4, 3.0, 2, 3, 1.0, 0.7, "rand"),
    ("n6_bimodal", 6, 4.0, 3, 6, 0.38, 1.0, "bimodal"),
    ("n8_bimodal", 8, 5.0, 4, 6, 0.38, 1.0, "bimodal"),
    ("n16_bkw", 16, 6.0, 8, 6, 0.0, 1.0 / (4 * math.pi), "bkw"),
]


def state(vm, kind, rng, npts=3):
    out = []
    for _ in range(npts):
        if kind == "rand":
            out.append(rng.random(vm.n))
        elif kind == "bkw":
            out.append(O.bkw(vm, 5.5))
        else:
            u1, u2 = rng.uniform(-0.6, 0.6, 3), rng.uniform(-0.6, 0.6, 3)
            t1, t2 = rng.uniform(0.5, 2.0, 2)
            out.append(0.5 * O.maxwellian(vm, 1, u1, t1) + 0.5 * O.maxwellian(vm, 1, u2, t2))
    return np.array(out)


def main():
    rng = np.random.default_rng(2024)
    for name, N, L, Nrho, M, gamma, prefac, kind in CASES:
        vm = O.VelocityMesh(N, L, Nrho, M)
        f = state(vm, kind, rng)
        Q = O.VHSOracle(vm, gamma, prefac).q(f)
        np.savez_compressed(os.path.join(HERE, f"vhs_{name}.npz"), f=f, Q=Q,
                            dims=np.array([N, Nrho, M]), params=np.array([L, gamma, prefac]))
        print(name, f.shape)


if __name__ == "__main__":
    main()
