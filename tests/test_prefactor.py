"""Configuration -> collision-kernel prefactor, pinned from the physics.

The independent check builds the dimensional VHS/VSS collision kernel from
Bird's variable-hard-sphere cross-section (Bird 1994, eqs. 4.63 and 4.64;
the VSS angular law of Koura & Matsumoto 1991) in SI units and rescales it
with the solver's reference scales (length H0, velocity
u0 = sqrt(2 R0 T0 / molarMass0), number density n0), then compares the
nondimensional kernel coefficient with:

* the oracle's restatements vhs_prefactor / vss_prefactor
  (frfs/solvers/dgfs/scattering.py:50-62, frfs/solvers/dgfsbi/scattering.py:58-92);
* the product's load_parameters through each of its configuration
  branches: dRef/Tref (the reference's keys), Kn, and a direct prefactor.

A wrong power of 2, a missing sqrt(2) or mass ratio, or a wrong exponent on
Tref/T0 or on the reduced mass changes the kernel by a constant factor and
fails here; the evaluation at several relative speeds and deflection angles
also fixes the exponents gamma = 2(1 - omega) and alpha - 1.  The hard-sphere
mean free path 1/(sqrt(2) pi n d^2) pins the Knudsen number itself.  CPU only.
"""
import math

import numpy as np
import pytest

from oracle import vhs as O
from oracle import vss as V

R0 = 8.3144598
NA = 6.0221409e23
KB = R0 / NA


def kernel_coefficient(omega, alpha, dRef, Tref, molar_p, molar_q, H0, T0, rho0, molar0):
    """Nondimensional B / (g^gamma (1 + cos chi)^(alpha - 1)) from the SI kernel.

    B(g, chi) = sigma_T(g) g alpha (1 + cos chi)^(alpha-1) / (2^(alpha+1) pi), the
    VSS law normalised so that its integral over the sphere is sigma_T g, with
    sigma_T = pi d^2 and Bird's VHS diameter
    d^2 = dRef^2 (2 kB Tref / (m_r g^2))^(omega - 1/2) / Gamma(5/2 - omega).
    Nondimensional time H0/u0, velocity u0 and density n0 give Bhat = B n0 H0 / u0.
    The coefficient is evaluated at several (g, chi) and must not depend on them."""
    mp, mq = molar_p / NA, molar_q / NA
    mr = mp * mq / (mp + mq)
    m0 = molar0 / NA
    n0 = rho0 / m0
    u0 = math.sqrt(2.0 * KB * T0 / m0)
    gamma = 2.0 * (1.0 - omega)
    vals = []
    for ghat, cosx in ((0.37, -0.6), (1.0, 0.0), (2.9, 0.8)):
        g = ghat * u0
        d2 = dRef ** 2 * (2.0 * KB * Tref / (mr * g * g)) ** (omega - 0.5) / math.gamma(2.5 - omega)
        sigma_t = math.pi * d2
        B = sigma_t * g * alpha * (1.0 + cosx) ** (alpha - 1.0) / (2.0 ** (alpha + 1.0) * math.pi)
        Bhat = B * n0 * H0 / u0
        vals.append(Bhat / (ghat ** gamma * (1.0 + cosx) ** (alpha - 1.0)))
    vals = np.array(vals)
    assert np.ptp(vals) <= 1e-12 * abs(vals[0])
    return float(vals[0])


SCALES = [dict(H0=1.0, T0=1.0, rho0=1.0, molar0=1.0),
          dict(H0=2.5e-4, T0=273.0, rho0=1.78e-4, molar0=39.948e-3)]
GASES = [dict(omega=0.5, dRef=1.0e-8, Tref=1.0), dict(omega=0.81, dRef=4.17e-10, Tref=273.0),
         dict(omega=1.0, dRef=3.0e-10, Tref=300.0)]


@pytest.mark.parametrize("sc", SCALES)
@pytest.mark.parametrize("gas", GASES)
def test_vhs_prefactor_is_birds_kernel(sc, gas):
    n0 = sc["rho0"] / sc["molar0"] * NA
    want = kernel_coefficient(gas["omega"], 1.0, gas["dRef"], gas["Tref"], sc["molar0"], sc["molar0"],
                              sc["H0"], sc["T0"], sc["rho0"], sc["molar0"])
    got = O.vhs_prefactor(gas["omega"], gas["dRef"], gas["Tref"], sc["H0"], sc["T0"], n0)
    assert got == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("masses", [(1.0, 1.0), (1.0, 2.1), (4.0, 1.0), (0.5, 39.9)])
@pytest.mark.parametrize("gas", GASES)
@pytest.mark.parametrize("alpha", [1.0, 1.4])
def test_vss_prefactor_is_birds_kernel(masses, gas, alpha):
    sc = SCALES[1]
    n0 = sc["rho0"] / sc["molar0"] * NA
    mp, mq = masses            # nondimensional: molar mass / molarMass0
    want = kernel_coefficient(gas["omega"], alpha, gas["dRef"], gas["Tref"], mp * sc["molar0"],
                              mq * sc["molar0"], sc["H0"], sc["T0"], sc["rho0"], sc["molar0"])
    got = V.vss_prefactor(mp, mq, gas["omega"], alpha, gas["dRef"], gas["Tref"], sc["H0"], sc["T0"], n0)
    assert got == pytest.approx(want, rel=1e-13)


def test_vss_reduces_to_vhs_for_one_species():
    n0 = NA
    for gas in GASES:
        a = V.vss_prefactor(1.0, 1.0, gas["omega"], 1.0, gas["dRef"], gas["Tref"], 1.0, 1.0, n0)
        b = O.vhs_prefactor(gas["omega"], gas["dRef"], gas["Tref"], 1.0, 1.0, n0)
        assert a == pytest.approx(b, rel=1e-14)


def test_bkw_kernel_from_knudsen():
    """Maxwell molecules (omega = 1) at invKn = Gamma(3/2): the kernel 1/(4 pi)
    for which the BKW solution holds (dK/dt = (1 - K)/6)."""
    from dgfs_hip.cases import bkw_cfg
    from dgfs_hip.solvers.dgfs.scattering import DGFSVHSGLLScatteringModel
    sm = _loaded(DGFSVHSGLLScatteringModel, bkw_cfg(), vhs=True)
    assert sm.prefactor == pytest.approx(1.0 / (4.0 * math.pi), rel=1e-15)
    assert sm._gamma == 0.0


def _loaded(cls, cfg, vhs):
    """A scattering model with parameters loaded and no device tables."""
    from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
    from dgfs_hip.solvers.dgfsbi.velocitymesh import DGFSBiVelocityMesh
    sm = object.__new__(cls)
    sm.cfg = cfg
    sm.vm = DGFSVelocityMesh(None, cfg) if vhs else DGFSBiVelocityMesh(None, cfg, 2)
    sm._nspcs = 1 if vhs else 2
    sm.load_parameters()
    return sm


def _vhs_cfg(sc, gas, key):
    from dgfs_hip.cases import base_cfg
    cfg = base_cfg(8, 5.0, 4, 6, omega=gas["omega"], Kn=1.0)
    for k, v in dict(H0=sc["H0"], T0=sc["T0"], rho0=sc["rho0"], molarMass0=sc["molar0"]).items():
        cfg.set("non-dim", k, v)
    cfg._cp.remove_option("scattering-model", "Kn")
    if key == "dRef":
        cfg.set("scattering-model", "dRef", gas["dRef"])
        cfg.set("scattering-model", "Tref", gas["Tref"])
    return cfg


@pytest.mark.parametrize("sc", SCALES)
@pytest.mark.parametrize("gas", GASES)
def test_product_vhs_branches(sc, gas):
    """dRef/Tref branch = Bird's kernel; the Kn branch at the Kn the dRef branch
    reports, and the direct prefactor branch, give the same kernel."""
    from dgfs_hip.solvers.dgfs.scattering import DGFSVHSGLLScatteringModel as C
    want = kernel_coefficient(gas["omega"], 1.0, gas["dRef"], gas["Tref"], sc["molar0"], sc["molar0"],
                              sc["H0"], sc["T0"], sc["rho0"], sc["molar0"])
    smd = _loaded(C, _vhs_cfg(sc, gas, "dRef"), vhs=True)
    assert smd.prefactor == pytest.approx(want, rel=1e-13)
    assert smd._gamma == pytest.approx(2.0 * (1.0 - gas["omega"]), abs=1e-15)
    cfg = _vhs_cfg(sc, gas, "Kn")
    cfg.set("scattering-model", "Kn", repr(smd._Kn))
    smk = _loaded(C, cfg, vhs=True)
    assert smk.prefactor == pytest.approx(smd.prefactor, rel=1e-14)
    cfg = _vhs_cfg(sc, gas, "prefactor")
    cfg.set("scattering-model", "prefactor", repr(want))
    assert _loaded(C, cfg, vhs=True).prefactor == want


@pytest.mark.parametrize("sc", SCALES)
def test_hard_sphere_knudsen_is_mean_free_path(sc):
    """omega = 1/2: Kn = lambda / H0 with lambda = 1 / (sqrt(2) pi n0 d^2)."""
    from dgfs_hip.solvers.dgfs.scattering import DGFSVHSGLLScatteringModel as C
    gas = dict(omega=0.5, dRef=3.7e-10, Tref=123.0)
    sm = _loaded(C, _vhs_cfg(sc, gas, "dRef"), vhs=True)
    n0 = sc["rho0"] / sc["molar0"] * NA
    lam = 1.0 / (math.sqrt(2.0) * math.pi * n0 * gas["dRef"] ** 2)
    assert sm._Kn == pytest.approx(lam / sc["H0"], rel=1e-14)


@pytest.mark.parametrize("alpha", [1.0, 1.3])
def test_product_vss_dref_branch(alpha):
    from dgfs_hip.cases import bi_couette_cfg
    from dgfs_hip.solvers.dgfsbi.scattering import DGFSBiVSSGLLScatteringModel as C
    sc = SCALES[1]
    masses = (1.0, 2.1)
    omega, dRef, Tref = 0.81, 4.17e-10, 273.0
    cfg = bi_couette_cfg(Nv=8, L=5.0, Nrho=4, masses=masses, omega=omega, alpha=alpha, dRef=dRef, Tref=Tref)
    for k, v in dict(H0=sc["H0"], T0=sc["T0"], rho0=sc["rho0"], molarMass0=sc["molar0"]).items():
        cfg.set("non-dim", k, v)
    for p, m in enumerate(masses):
        cfg.set("velocity-mesh", f"molarMass{p + 1}", m * sc["molar0"])
    sm = _loaded(C, cfg, vhs=False)
    for p in range(2):
        for q in range(2):
            want = kernel_coefficient(omega, alpha, dRef, Tref, masses[p] * sc["molar0"],
                                      masses[q] * sc["molar0"], sc["H0"], sc["T0"], sc["rho0"], sc["molar0"])
            assert sm.prefactor[p, q] == pytest.approx(want, rel=1e-13)
            assert sm._eta[p, q] == pytest.approx(alpha - 1.0, abs=1e-15)
