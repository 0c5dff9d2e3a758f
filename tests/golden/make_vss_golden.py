"""Writes tests/golden/vss_*.npz: per-species (f, Q) of the FP64 multi-species
VSS oracle at small grids, used as fixed vectors by the GPU parity tests.
Run from the repo root:
    python tests/golden/make_vss_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import vss as V  # noqa: E402

PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]
CASES = [
    # name, N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac, kind
    ("n4_rand", 4, 5.0, 2, 6, 12, (1.0, 2.1), 1.3, (0.81, 0.805, 0.805, 0.80), (0.7, 1.1, 0.9, 1.3), "rand"),
    ("n8_bimodal", 8, 5.0, 4, 12, 24, (1.0, 2.1), 1.0, (0.81, 0.805, 0.805, 0.80), (1.0, 1.0, 1.0, 1.0), "bimodal"),
    ("n12_vss", 12, 6.0, 6, 12, 24, (1.0, 2.1), 1.4, (0.75, 0.77, 0.77, 0.79), (1.0, 0.8, 0.8, 0.6), "bimodal"),
]


def states(vm, kind, rng, npts=3):
    out = [[], []]
    for _ in range(npts):
        for p in range(2):
            if kind == "rand":
                out[p].append(rng.random(vm.n))
            else:
                u1, u2 = rng.uniform(-0.5, 0.5, 3), rng.uniform(-0.5, 0.5, 3)
                t1, t2 = rng.uniform(0.6, 1.6, 2)
                out[p].append(0.5 * V.species_maxwellian(vm, p, 1.0, u1, t1)
                              + 0.5 * V.species_maxwellian(vm, p, 0.7, u2, t2))
    return [np.array(o) for o in out]


def main():
    rng = np.random.default_rng(77)
    for name, N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac, kind in CASES:
        vm = V.BiVelocityMesh(N, L, Nrho, M, masses)
        g = {k: 2.0 * (1.0 - w) for k, w in zip(PAIRS, omegas)}
        eta = {k: alpha - 1.0 for k in PAIRS}
        pf = dict(zip(PAIRS, prefac))
        orc = V.VSSOracle(vm, g, eta, pf, V.full_sphere_design(Mpre))
        f0, f1 = states(vm, kind, rng)
        Q0, Q1 = orc.q([f0, f1])
        np.savez_compressed(os.path.join(HERE, f"vss_{name}.npz"), f0=f0, f1=f1, Q0=Q0, Q1=Q1,
                            dims=np.array([N, Nrho, M, Mpre]), L=L, masses=np.array(masses),
                            alpha=alpha, omegas=np.array(omegas), prefac=np.array(prefac))
        print(name, f0.shape)


if __name__ == "__main__":
    main()
