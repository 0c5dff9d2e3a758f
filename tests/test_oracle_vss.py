"""Pins of the multi-species VSS oracle (oracle/vss.py).  CPU only.

The reference has no tests or golden vectors for this operator; the oracle
is pinned by an explicit-DFT evaluation, by its reduction to the
single-species VHS oracle (itself pinned by the BKW closed form), by the
equilibrium of species Maxwellians and by measured conservation."""
import math

import numpy as np
import pytest

from oracle import vhs as H
from oracle import vss as V

PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def _params(omegas=(0.81, 0.805, 0.805, 0.80), alpha=1.0):
    g = {k: 2.0 * (1.0 - w) for k, w in zip(PAIRS, omegas)}
    eta = {k: alpha - 1.0 for k in PAIRS}
    pf = {k: 1.0 for k in PAIRS}
    return g, eta, pf


@pytest.mark.parametrize("alpha", [1.0, 1.3])
def test_matches_bruteforce_dft(alpha):
    vm = V.BiVelocityMesh(4, 5.0, 2, 6, (1.0, 2.1))
    g, eta, pf = _params(alpha=alpha)
    orc = V.VSSOracle(vm, g, eta, pf, V.full_sphere_design(12))
    rng = np.random.default_rng(3)
    fs = [rng.random(64), rng.random(64)]
    for p, q in PAIRS:
        a = orc.qij(p, q, fs[p], fs[q])
        b = V.q_bruteforce(vm, orc, p, q, fs[p], fs[q])
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-13


def test_reduces_to_single_species_vhs():
    """Equal masses, alpha = 1 and the exact spherical average give the VHS
    operator: the full-sphere sum of t3 t4 equals twice the half-sphere sum of
    P+ P-, and b, c become the VHS b1, b2."""
    N, L, Nrho, Mh = 8, 5.0, 4, 6
    vm = V.BiVelocityMesh(N, L, Nrho, 2 * Mh, (1.0, 1.0))
    g = {k: 0.38 for k in PAIRS}
    eta = {k: 0.0 for k in PAIRS}
    pf = {k: 1.0 for k in PAIRS}
    orc = V.VSSOracle(vm, g, eta, pf, exact_sphere=True)
    hv = H.VelocityMesh(N, L, Nrho, Mh)
    ho = H.VHSOracle(hv, 0.38, 1.0)
    f = H.maxwellian(hv, 1.0, [0.3, 0, 0], 1.0) + 0.5 * H.maxwellian(hv, 1.0, [-0.5, 0.2, 0], 0.7)
    a = orc.qij(0, 0, f, f)
    b = ho.q(f)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-13


def test_design_tables_approach_exact_average():
    """b built from a 24-point design matches the exact spherical average
    (4 pi sinc) at low frequencies, where a 5-design integrates the plane
    wave's Taylor terms exactly up to degree 5."""
    vm = V.BiVelocityMesh(8, 5.0, 4, 12, (1.0, 2.1))
    bd, cd = V.vss_tables(vm, 1.0, 2.1, 0.38, 0.0, V.full_sphere_design(24))
    be, ce = V.vss_tables(vm, 1.0, 2.1, 0.38, 0.0, exact_sphere=True)
    l = H.lvectors(8)
    low = np.sqrt((l ** 2).sum(0)) <= 1.0
    assert np.max(np.abs(bd[:, low] - be[:, low])) / np.max(np.abs(be)) < 2e-3
    assert np.max(np.abs(cd[low] - ce[low])) / np.max(np.abs(ce)) < 2e-3


def test_species_maxwellians_are_equilibrium():
    """Q_p(M_0, M_1) -> 0 for Maxwellians with a common velocity and
    temperature, converging with the grid (spectral accuracy of the method)."""
    g, eta, pf = _params()
    res = []
    for N, L in ((16, 5.0), (24, 6.0)):
        vm = V.BiVelocityMesh(N, L, 8, 12, (1.0, 2.1))
        orc = V.VSSOracle(vm, g, eta, pf, exact_sphere=True)
        M0 = V.species_maxwellian(vm, 0, 1.0, [0.2, 0, 0], 1.0)
        M1 = V.species_maxwellian(vm, 1, 0.5, [0.2, 0, 0], 1.0)
        B0 = 0.5 * V.species_maxwellian(vm, 0, 1.0, [0.6, 0, 0], 0.8) + \
            0.5 * V.species_maxwellian(vm, 0, 1.0, [-0.6, 0.1, 0], 1.1)
        B1 = V.species_maxwellian(vm, 1, 0.5, [0, 0.3, 0], 0.7)
        Qe, Qb = orc.q([M0, M1]), orc.q([B0, B1])
        res.append([np.linalg.norm(Qe[p]) / np.linalg.norm(Qb[p]) for p in range(2)])
    # measured: [0.019, 0.113] at 16^3, [0.0026, 0.020] at 24^3
    assert res[1][0] < 4e-3 and res[1][1] < 3e-2
    assert res[0][0] > 4 * res[1][0] and res[0][1] > 4 * res[1][1]


def test_conservation_is_measured():
    """Species mass per pair and the pair's momentum / energy exchange are
    conserved to the discretisation error of the spectral method (measured,
    a few percent of int |Q| at 24^3; the FP64 projection of the product
    removes it)."""
    g, eta, pf = _params()
    vm = V.BiVelocityMesh(24, 6.0, 8, 12, (1.0, 2.1))
    orc = V.VSSOracle(vm, g, eta, pf, V.full_sphere_design(24))
    B0 = 0.5 * V.species_maxwellian(vm, 0, 1.0, [0.6, 0, 0], 0.8) + \
        0.5 * V.species_maxwellian(vm, 0, 1.0, [-0.6, 0.1, 0], 1.1)
    B1 = V.species_maxwellian(vm, 1, 0.5, [0, 0.3, 0], 0.7)
    cv, cw, m = vm.cv, vm.cw, vm.masses
    Q01, Q10 = orc.qij(0, 1, B0, B1), orc.qij(1, 0, B1, B0)
    sc = np.abs(Q01).sum() * cw
    assert abs(Q01.sum() * cw) / sc < 1e-2
    mom = (m[0] * (cv * Q01).sum(1) + m[1] * (cv * Q10).sum(1)) * cw / sc
    en = (m[0] * ((cv ** 2).sum(0) * Q01).sum() + m[1] * ((cv ** 2).sum(0) * Q10).sum()) * cw / sc
    assert np.max(np.abs(mom)) < 3e-2 and abs(en) < 3e-2


def test_bimolecular_velocity_mesh_support():
    """S and R follow the mass-ratio rule of dgfsbi/velocitymesh.py:72-76 and
    reduce to the single-species 2L/(3+sqrt 2) for equal masses."""
    vm = V.BiVelocityMesh(8, 6.0, 4, 12, (1.0, 1.0))
    assert abs(vm.S - 2 * 6.0 / (3 + math.sqrt(2))) < 1e-14
    vm2 = V.BiVelocityMesh(8, 6.0, 4, 12, (1.0, 2.1))
    mr = 2.1
    s1 = 12.0 / (max(4 * mr / (1 + mr), 2) + math.sqrt(1 + mr) + 1)
    s2 = 12.0 / (max(4 / (1 + mr), 2) + math.sqrt(1 + 1 / mr) + 1)
    assert abs(vm2.S - min(s1, s2)) < 1e-14 and vm2.S < vm.S
