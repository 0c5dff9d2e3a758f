"""Configuration builders for the benchmark workloads (INI keys of the
reference, frfs/inifile.py; nondimensional units H0 = T0 = rho0 =
molarMass0 = 1 so lengths are in H0 and velocities in u0 = sqrt(2 R0))."""
import math

from dgfs_hip.inifile import Inifile

U0 = math.sqrt(2 * 8.3144598)


def base_cfg(Nv, L, Nrho, M, omega=0.5, Kn=None, prefactor=None, order=2, NvBatchSize=None,
             precision="double", project=None):
    cfg = Inifile()
    for k, v in dict(H0=1.0, T0=1.0, rho0=1.0, molarMass0=1.0).items():
        cfg.set("non-dim", k, v)
    cfg.set("constants", "Nv", Nv)
    cfg.set("constants", "NvBatchSize", NvBatchSize or Nv ** 3)
    cfg.set("constants", "Nrho", Nrho)
    cfg.set("velocity-mesh", "cmax", 0.0)
    cfg.set("velocity-mesh", "Tmax", 1.0)
    cfg.set("velocity-mesh", "dev", L)
    cfg.set("spherical-design-rule", "ssrule", "womersley")
    cfg.set("spherical-design-rule", "M", M)
    cfg.set("scattering-model", "type", "vhs-gll")
    cfg.set("scattering-model", "omega", omega)
    if prefactor is not None:
        cfg.set("scattering-model", "prefactor", prefactor)
    else:
        cfg.set("scattering-model", "Kn", Kn if Kn is not None else 1.0)
    cfg.set("scattering-model", "precision", precision)
    if project is not None:
        cfg.set("scattering-model", "project", "true" if project else "false")
    cfg.set("solver", "system", "dgfs")
    cfg.set("solver", "order", order)
    cfg.set("solver-interfaces", "riemann-solver", "rusanov")
    cfg.set("solver-interfaces-line", "flux-pts", "gauss-legendre")
    cfg.set("solver-elements-quad", "soln-pts", "gauss-legendre")
    cfg.set("solver-elements-tri", "soln-pts", "williams-shunn")
    return cfg


def maxwellian_section(cfg, sect, rho, T, u, kind=None):
    if kind:
        cfg.set(sect, "type", kind)
    cfg.set(sect, "rho", rho)
    cfg.set(sect, "T", T)
    cfg.set(sect, "ux", u[0] * U0)
    cfg.set(sect, "uy", u[1] * U0)
    cfg.set(sect, "uz", u[2] * U0)


def couette_cfg(Nv=24, L=7.0, Nrho=12, M=12, Kn=0.1, omega=0.5, order=2, uwall=0.5, **kw):
    """Planar Couette flow: x-periodic channel, diffuse walls at T=1 moving at
    +-uwall, equilibrium Maxwellian initial state (rho 1, T 1)."""
    cfg = base_cfg(Nv, L, Nrho, M, omega=omega, Kn=Kn, order=order, **kw)
    maxwellian_section(cfg, "soln-ics", 1.0, 1.0, (0, 0, 0), "maxwellian")
    maxwellian_section(cfg, "soln-bcs-bottom", 1.0, 1.0, (-uwall, 0, 0), "dgfs-wall-diffuse")
    maxwellian_section(cfg, "soln-bcs-top", 1.0, 1.0, (uwall, 0, 0), "dgfs-wall-diffuse")
    return cfg


def cavity_cfg(Nv=32, L=6.0, Nrho=16, M=12, Kn=0.05, omega=0.81, order=3, ulid=0.3, **kw):
    """Lid-driven cavity: diffuse walls at T=1, the top one moving at ulid."""
    cfg = base_cfg(Nv, L, Nrho, M, omega=omega, Kn=Kn, order=order, **kw)
    maxwellian_section(cfg, "soln-ics", 1.0, 1.0, (0, 0, 0), "maxwellian")
    for name in ("bottom", "left", "right"):
        maxwellian_section(cfg, f"soln-bcs-{name}", 1.0, 1.0, (0, 0, 0), "dgfs-wall-diffuse")
    maxwellian_section(cfg, "soln-bcs-top", 1.0, 1.0, (ulid, 0, 0), "dgfs-wall-diffuse")
    return cfg


def bkw_cfg(Nv=16, L=6.0, Nrho=8, M=6, npts=4, t0=5.5, omega=1.0, Kn=None, precision="double",
            project=None, NvBatchSize=None):
    """Spatially homogeneous BKW relaxation of Maxwell molecules: `npts`
    points at the BKW state of time t0 (frfs/solvers/dgfs/initcond.py:120-140).
    The BKW solution holds for the kernel B = 1/(4 pi) (omega = 1), which the
    prefactor formula gives at invKn = Gamma(3/2), i.e. Kn = 2/sqrt(pi)."""
    if Kn is None:
        Kn = 1.0 / math.gamma(1.5)
    cfg = base_cfg(Nv, L, Nrho, M, omega=omega, Kn=Kn, NvBatchSize=NvBatchSize, precision=precision,
                   project=project)
    cfg.set("soln-ics", "type", "bkw")
    cfg.set("soln-ics", "t0", t0)
    cfg.set("mesh", "points", npts)
    return cfg


def _bi_base(Nv, L, Nrho, M, Mpre, masses, omega, alpha, prefactor, dRef, Tref, order, NvBatchSize,
             precision, project):
    """Common keys of the multi-species configurations (frfs/solvers/dgfsbi/*).
    omega / alpha / prefactor: one value for every pair or a dict keyed by
    (p, q)."""
    cfg = Inifile()
    for k, v in dict(H0=1.0, T0=1.0, rho0=1.0, molarMass0=1.0).items():
        cfg.set("non-dim", k, v)
    cfg.set("constants", "Nv", Nv)
    cfg.set("constants", "NvBatchSize", NvBatchSize or Nv ** 3)
    cfg.set("constants", "Nrho", Nrho)
    cfg.set("velocity-mesh", "cmax", 0.0)
    cfg.set("velocity-mesh", "Tmax", 1.0)
    cfg.set("velocity-mesh", "dev", L)
    for p, m in enumerate(masses):
        cfg.set("velocity-mesh", f"molarMass{p + 1}", m)
    cfg.set("spherical-design-rule", "ssrule", "womersley")
    cfg.set("spherical-design-rule", "M", M)
    cfg.set("scattering-model", "type", "vss-gll")
    cfg.set("scattering-model", "Mpre", Mpre)
    cfg.set("scattering-model", "precision", precision)
    if project is not None:
        cfg.set("scattering-model", "project", "true" if project else "false")
    S = len(masses)
    val = lambda v, p, q: v[(p, q)] if isinstance(v, dict) else v
    for p in range(S):
        for q in range(S):
            pq = f"{p}{q}"
            cfg.set("scattering-model", "alpha" + pq, val(alpha, p, q))
            cfg.set("scattering-model", "omega" + pq, val(omega, p, q))
            if prefactor is not None:
                cfg.set("scattering-model", "prefactor" + pq, val(prefactor, p, q))
            else:
                cfg.set("scattering-model", "dRef" + pq, dRef)
                cfg.set("scattering-model", "Tref" + pq, Tref)
    cfg.set("solver", "system", "dgfsbi")
    cfg.set("solver", "order", order)
    cfg.set("solver-interfaces", "riemann-solver", "rusanov")
    cfg.set("solver-interfaces-line", "flux-pts", "gauss-legendre")
    cfg.set("solver-elements-quad", "soln-pts", "gauss-legendre")
    cfg.set("solver-elements-tri", "soln-pts", "williams-shunn")
    return cfg


N0 = 6.0221409e+23   # n0 = rho0 / molarMass0 * N_A with rho0 = molarMass0 = 1


def _bi_state(cfg, sect, T, u, nden=None, kind=None):
    if kind:
        cfg.set(sect, "type", kind)
    cfg.set(sect, "T", T)
    cfg.set(sect, "ux", u[0] * U0)
    cfg.set(sect, "uy", u[1] * U0)
    cfg.set(sect, "uz", u[2] * U0)
    for p, nd in enumerate(nden or ()):
        cfg.set(sect, f"nden{p + 1}", nd * N0)


def bi_couette_cfg(Nv=16, L=5.0, Nrho=4, M=12, Mpre=24, masses=(1.0, 2.1), nden=(0.5, 0.5),
                   omega=0.5, alpha=1.0, prefactor=None, dRef=1.0, Tref=1.0, order=2, uwall=0.3,
                   NvBatchSize=None, precision="double", project=None):
    """Two-species (or more) planar Couette flow: x-periodic channel between
    diffuse walls at T = 1 moving at +-uwall, initial species Maxwellians at
    rest with number densities nden (in n0); VSS scattering with the same
    omega and alpha for every pair (keys of frfs/solvers/dgfsbi/*).  With
    `prefactor` set, every pair uses that kernel prefactor directly."""
    cfg = _bi_base(Nv, L, Nrho, M, Mpre, masses, omega, alpha, prefactor, dRef, Tref, order, NvBatchSize,
                   precision, project)
    cfg.set("soln-ics", "type", "maxwellian")
    _bi_state(cfg, "soln-ics", 1.0, (0.0, 0, 0), nden)
    _bi_state(cfg, "soln-bcs-bottom", 1.0, (-uwall, 0, 0), kind="dgfsbi-wall-diffuse")
    _bi_state(cfg, "soln-bcs-top", 1.0, (uwall, 0, 0), kind="dgfsbi-wall-diffuse")
    return cfg


def vhs_prefactor(Kn, omega):
    """Kernel prefactor of a VHS pair at Knudsen number Kn (invKn = 1/Kn in
    the prefactor of frfs/solvers/dgfs/scattering.py:50-65, alpha = 1)."""
    return (1.0 / Kn) / (2.0 ** (3.0 - omega) * math.gamma(2.5 - omega) * math.pi)


def cylinder_bi_cfg(Nv=24, L=6.0, Nrho=12, M=12, Mpre=24, masses=(1.0, 2.1), nden=(0.5, 0.5),
                    omegas=(0.81, 0.805, 0.80), Kn=0.05, mach=2.0, Twall=1.0, order=2, NvBatchSize=None,
                    precision="single", project=True):
    """Two-species VHS mixture (alpha = 1) in supersonic flow along +x past a
    cylinder: free-stream species Maxwellians (number densities nden, T = 1)
    at Mach `mach` of the mixture (monatomic, number-weighted mass) on the
    inlet and as the initial state, extrapolation on the outlet, and a
    diffuse wall at rest at Twall on the cylinder.  omegas = (w00, w01 = w10,
    w11); every pair's prefactor is the VHS one at Kn."""
    w = {(0, 0): omegas[0], (0, 1): omegas[1], (1, 0): omegas[1], (1, 1): omegas[2]}
    pref = {k: vhs_prefactor(Kn, v) for k, v in w.items()}
    cfg = _bi_base(Nv, L, Nrho, M, Mpre, masses, w, 1.0, pref, None, None, order, NvBatchSize, precision,
                   project)
    mmix = sum(n * m for n, m in zip(nden, masses)) / sum(nden)
    uinf = mach * math.sqrt(5.0 / 3.0 / (2.0 * mmix))          # in u0 = sqrt(2 R T0 / M0)
    cfg.set("soln-ics", "type", "maxwellian")
    _bi_state(cfg, "soln-ics", 1.0, (uinf, 0, 0), nden)
    _bi_state(cfg, "soln-bcs-inlet", 1.0, (uinf, 0, 0), nden, kind="dgfsbi-inlet")
    cfg.set("soln-bcs-outlet", "type", "dgfsbi-outlet")
    cfg.set("soln-bcs-cylinder", "type", "dgfsbi-wall-diffuse-cyl")
    for k in ("ux", "uy", "uz"):
        cfg.set("soln-bcs-cylinder", k, "0")
    cfg.set("soln-bcs-cylinder", "T", Twall)
    return cfg
