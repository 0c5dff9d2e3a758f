"""GPU parity of the multi-species DGFS system (dgfsbi) against the FP64
oracles: FR advection of every species (oracle/fr.py, diffuse walls with
each species' unit-density wall Maxwellian) plus the VSS collision operator
Q_p = sum_q Q_pq (oracle/vss.py), stepped with the reference's dgfsbi-euler
and dgfsbi-tvd-rk2 sequences.  FP64 bar: 1e-11 relative L2 per species."""
import sys

import numpy as np
import pytest

from oracle import fr as OF
from oracle import vss as V

pytestmark = pytest.mark.gpu
MASSES = (1.0, 2.1)
PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# 16^3 on [-4, 4]^3: the finest grid at which the heavier species' Maxwellian
# passes the initial condition's 1e-5 moment check (h = 0.5)
NV, LV, ORDER = 16, 4.0, 1
NUPTS = (ORDER + 1) ** 2


def build(nregs, NvBS=None, Nv=NV):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import bi_couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    cfg = bi_couette_cfg(Nv=Nv, L=LV, Nrho=4, M=12, Mpre=24, masses=MASSES, nden=(0.6, 0.4),
                         omega=0.5, alpha=1.2, prefactor=0.05, order=ORDER, uwall=0.3, NvBatchSize=NvBS)
    be = HIPBackend(cfg)
    mesh = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    return be, DGFSBiSystem(be, mesh, cfg, nregs)


def oracles(sysm):
    vm = sysm.vm
    ov = V.BiVelocityMesh(vm.Nv(), vm.L(), vm.Nrho(), vm.M(), MASSES)
    assert np.allclose(ov.cv, vm.cv()) and abs(ov.cw - vm.cw()) < 1e-15
    omesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True)
    adv = []
    for p in range(2):
        walls = {"bottom": (np.array([-0.3, 0, 0]), V.species_maxwellian(ov, p, 1.0, [-0.3, 0, 0], 1.0)),
                 "top": (np.array([0.3, 0, 0]), V.species_maxwellian(ov, p, 1.0, [0.3, 0, 0], 1.0))}
        adv.append(OF.AdvectionOracle(ORDER, omesh, ov.cv, ov.cw, walls))
    g = {k: 2.0 * (1.0 - 0.5) for k in PAIRS}
    eta = {k: 0.2 for k in PAIRS}
    qo = V.VSSOracle(ov, g, eta, {k: 0.05 for k in PAIRS}, V.full_sphere_design(24))

    def Q(fs):
        fl = [f.transpose(2, 0, 1).reshape(-1, ov.n) for f in fs]
        return [q.reshape(6, NUPTS, ov.n).transpose(1, 2, 0) for q in qo.q(fl)]
    return ov, adv, Q


def initial(ov, seed):
    rng = np.random.default_rng(seed)
    return [V.species_maxwellian(ov, p, nd, [0.05, -0.02, 0], 1.0)[None, :, None]
            * (1 + 0.1 * rng.random((NUPTS, 1, 6))) for p, nd in enumerate((0.6, 0.4))]


def test_bi_initial_condition_and_registers(device):
    be, sysm = build(6)
    ov, _, _ = oracles(sysm)
    say("bi registers", len(sysm.eles_scal_upts_full))
    assert len(sysm.eles_scal_upts_full) == 6
    for p, nd in enumerate((0.6, 0.4)):
        f = sysm.species_register(p, 0).get()
        want = V.species_maxwellian(ov, p, nd, [0, 0, 0], 1.0)
        assert np.allclose(f, want[None, :, None], rtol=1e-14, atol=0)


@pytest.mark.parametrize("stepper,nbs", [("dgfsbi-euler", None), ("dgfsbi-tvd-rk2", None),
                                         ("dgfsbi-tvd-rk2", 1024)])
def test_bi_step_vs_oracle(stepper, nbs, device):
    from dgfs_hip.integrators.steppers import get_stepper
    n0 = 2 if stepper == "dgfsbi-euler" else 3
    be, sysm = build(2 * n0, NvBS=nbs)
    ov, adv, Q = oracles(sysm)
    f0 = initial(ov, 11)
    for p in range(2):
        sysm.species_register(p, 0).set(f0[p])
    dt = 0.01
    st = get_stepper(stepper, be, sysm, dt)
    say("bi step", stepper, "NvBatchSize", sysm.vm.NvBatchSize())
    st.advance_to(2 * dt)
    be.synchronize()
    got = [sysm.species_register(p, 0).get() for p in range(2)]
    f = [x.copy() for x in f0]
    for _ in range(2):
        q0 = Q(f)
        f1 = [f[p] + dt * (adv[p].rhs(f[p]) + q0[p]) for p in range(2)]
        if stepper == "dgfsbi-euler":
            f = f1
        else:
            q1 = Q(f1)
            f = [0.5 * f[p] + 0.5 * f1[p] + 0.5 * dt * (adv[p].rhs(f1[p]) + q1[p]) for p in range(2)]
    for p in range(2):
        err = rel(got[p], f[p])
        say("  species", p, "rel L2", err)
        assert err < 1e-11


def test_bi_species_number_conserved(device):
    """Resting diffuse walls reflect every incoming particle of each species
    and the FP64 projection keeps every species' number in Q: the
    quadrature-weighted number of each species is conserved by the steps
    (1e-12), and a mixture of species Maxwellians at the walls' state stays
    put up to the operator's equilibrium error."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import bi_couette_cfg
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    cfg = bi_couette_cfg(Nv=NV, L=LV, Nrho=4, M=12, Mpre=24, masses=MASSES, nden=(0.6, 0.4),
                         omega=0.5, alpha=1.2, prefactor=0.05, order=ORDER, uwall=0.0, project=True)
    be = HIPBackend(cfg)
    mesh = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    sysm = DGFSBiSystem(be, mesh, cfg, 6)
    W = OF.QuadFR(ORDER).W
    ov = V.BiVelocityMesh(NV, LV, 4, 12, MASSES)

    def number(f):
        return float(np.einsum("i,ive->", W, f))

    eq = [sysm.species_register(p, 0).get() for p in range(2)]
    f0 = initial(ov, 5)
    for p in range(2):
        sysm.species_register(p, 0).set(f0[p])
    n0 = [number(f) for f in f0]
    st = get_stepper("dgfsbi-tvd-rk2", be, sysm, 0.01)
    say("bi number conservation, 3 steps")
    st.advance_to(0.03)
    be.synchronize()
    for p in range(2):
        f = sysm.species_register(p, 0).get()
        d = abs(number(f) - n0[p]) / n0[p]
        say("  species", p, "number drift", d)
        assert d < 1e-12
    for p in range(2):
        sysm.species_register(p, 0).set(eq[p])
    st2 = get_stepper("dgfsbi-tvd-rk2", be, sysm, 0.01)
    st2.advance_to(0.03)
    be.synchronize()
    for p in range(2):
        f = sysm.species_register(p, 0).get()
        ch = rel(f, eq[p])
        say("  species", p, "equilibrium change", ch)
        assert ch < 1e-3


def test_bi_cyl_wall_vs_oracle(device):
    """dgfsbi-wall-diffuse-cyl (inters.py:679-763): per-point wall velocity
    and temperature from expressions of x, each species emitting
    exp(-m_p |c - U|^2 / T), against the oracle's per-point diffuse wall
    at temperature T / m_p (1e-12 per species)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import U0, bi_couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    cfg = bi_couette_cfg(Nv=NV, L=LV, Nrho=4, M=12, Mpre=24, masses=MASSES, nden=(0.6, 0.4),
                         omega=0.5, alpha=1.2, prefactor=0.05, order=ORDER, uwall=0.3)
    cfg.set("constants", "uw", 0.2 * U0)
    sect = "soln-bcs-bottom"
    cfg.set(sect, "type", "dgfsbi-wall-diffuse-cyl")
    cfg.set(sect, "ux", "uw*cos(2*pi*x/1.5)")
    cfg.set(sect, "uy", "0")
    cfg.set(sect, "uz", "0")
    cfg.set(sect, "T", "1.1 - 0.1*x")
    be = HIPBackend(cfg)
    sysm = DGFSBiSystem(be, rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True), cfg, 4)
    ov = V.BiVelocityMesh(NV, LV, 4, 12, MASSES)
    omesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True)
    f0 = initial(ov, 13)
    for p in range(2):
        m = MASSES[p]

        def wall(xy, m=m):
            U = np.zeros((len(xy), 3))
            U[:, 0] = 0.2 * np.cos(2 * np.pi * xy[:, 0] / 1.5)
            return U, (1.1 - 0.1 * xy[:, 0]) / m

        top = (np.array([0.3, 0, 0]), V.species_maxwellian(ov, p, 1.0, [0.3, 0, 0], 1.0))
        orc = OF.AdvectionOracle(ORDER, omesh, ov.cv, ov.cw, {"bottom": wall, "top": top})
        sysm.ele.scal_upts[0].set(f0[p])
        sysm.updatebc(0.0, p, 0, 1, 0, 0)
        sysm.rhs(0.0, p, 0, 1, 0, 0)
        be.synchronize()
        err = rel(sysm.ele.scal_upts[1].get(), orc.rhs(f0[p]))
        say("bi cyl wall species", p, "rel L2", err)
        assert err < 1e-12


def test_bi_fp32_steps_track_fp64(device):
    """The two-species system with the FP32 operator (register-resident gain
    at 24^3, FP64 mixture projection) against the FP64 operator over two
    dgfsbi-tvd-rk2 steps of the Couette channel at 24^3: per-species
    relative L2 within 1e-5, and each species' number conserved to 1e-12
    (resting walls)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import bi_couette_cfg
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    out = {}
    for prec, proj in (("single", True), ("double", True)):
        cfg = bi_couette_cfg(Nv=24, L=6.0, Nrho=6, M=12, Mpre=24, masses=MASSES, nden=(0.6, 0.4),
                             omega=0.5, alpha=1.0, prefactor=0.05, order=ORDER, uwall=0.0,
                             precision=prec, project=proj)
        be = HIPBackend(cfg)
        sysm = DGFSBiSystem(be, rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True), cfg, 6)
        ov = V.BiVelocityMesh(24, 6.0, 6, 12, MASSES)
        f0 = initial(ov, 17)
        for p in range(2):
            sysm.species_register(p, 0).set(f0[p])
        st = get_stepper("dgfsbi-tvd-rk2", be, sysm, 0.005)
        say("bi", prec, "2 steps at 24^3")
        st.advance_to(0.01)
        be.synchronize()
        if prec == "single":
            assert sysm.sm.op.get_algorithm() == 7   # alpha = 1: the isotropic slab gain
        out[prec] = [sysm.species_register(p, 0).get() for p in range(2)]
    W = OF.QuadFR(ORDER).W
    for p in range(2):
        err = rel(out["single"][p], out["double"][p])
        n0 = float(np.einsum("i,ive->", W, f0[p]))
        n1 = float(np.einsum("i,ive->", W, out["single"][p]))
        say("  species", p, "fp32 vs fp64", err, "number drift", abs(n1 - n0) / n0)
        assert err < 1e-5
        assert abs(n1 - n0) <= 1e-12 * n0


def test_bi_residual_plugin(device):
    """DGFSResidualBiPlugin (plugins/dgfsresidualbi.py:9-95): every nsteps
    steps, per species sqrt(sum (f_n - f_{n-1})^2 / sum f_{n-1}^2), against
    numpy on the unpacked registers (1e-12)."""
    import math
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.plugins import DGFSResidualBiPlugin
    be, sysm = build(6)
    ov, _, _ = oracles(sysm)
    f0 = initial(ov, 23)
    for p in range(2):
        sysm.species_register(p, 0).set(f0[p])
    st = get_stepper("dgfsbi-tvd-rk2", be, sysm, 0.01)
    seen = []
    st.plugins.append(lambda s: seen.append([sysm.species_register(p, 0).get() for p in range(2)]))
    pl = DGFSResidualBiPlugin(st, 2)
    st.plugins.append(pl)
    say("bi residual plugin, 4 steps")
    st.advance_to(0.04)
    assert len(pl.rows) == 2 and all(len(r) == 3 for r in pl.rows)
    for k, (i, j) in enumerate([(0, 1), (2, 3)]):
        for p in range(2):
            want = math.sqrt(np.sum((seen[j][p] - seen[i][p]) ** 2) / np.sum(seen[i][p] ** 2))
            say("  ", k, p, pl.rows[k][1 + p], want)
            assert abs(pl.rows[k][1 + p] - want) <= 1e-12 * want


def _writer_bi_2d(fs, cv, cw, masses, rho0, u0, T0, M0, R0, NA):
    """plugins/dgfsmomwriterbi.py:133-256 on host arrays f_p (nupts, vsize, neles)."""
    out, tot, mom = [], 0.0, 0.0
    sp = []
    for p, f in enumerate(fs):
        mr = masses[p]
        rho = f.sum(1) * mr * cw
        ux = np.tensordot(f, cv[0], axes=(1, 0)) * mr * cw / rho
        uy = np.tensordot(f, cv[1], axes=(1, 0)) * mr * cw / rho
        cx = cv[0][None, :, None] - ux[:, None, :]
        cy = cv[1][None, :, None] - uy[:, None, :]
        cz = cv[2][None, :, None]
        T = (f * (cx * cx + cy * cy + cz * cz)).sum(1) * (2.0 / 3.0 * mr * cw * mr) / rho
        sp.append((rho, ux, uy, T))
        tot = tot + rho
        mom = mom + rho * np.stack([ux, uy])
    W = mom / tot
    for p, f in enumerate(fs):
        mr = masses[p]
        rho, ux, uy, T = sp[p]
        cx = cv[0][None, :, None] - W[0][:, None, :]
        cy = cv[1][None, :, None] - W[1][:, None, :]
        cz = cv[2][None, :, None]
        c2 = cx * cx + cy * cy + cz * cz
        d = {"rho": rho * rho0, "Ux": ux * u0, "Uy": uy * u0, "T": T * T0,
             "Qx": mr * (f * c2 * cx).sum(1) * mr * cw * 0.5 * rho0 * u0 ** 3,
             "Qy": mr * (f * c2 * cy).sum(1) * mr * cw * 0.5 * rho0 * u0 ** 3,
             "Pxx": 2 * mr * (f * cx * cx).sum(1) * mr * cw * 0.5 * rho0 * u0 ** 2,
             "Pxy": 2 * mr * (f * cx * cy).sum(1) * mr * cw * 0.5 * rho0 * u0 ** 2,
             "Pyy": 2 * mr * (f * cy * cy).sum(1) * mr * cw * 0.5 * rho0 * u0 ** 2}
        d["p"] = (mr * R0 / M0) * d["rho"] * d["T"]
        d["nden"] = (NA / mr / M0) * d["rho"]
        out.append(d)
    return out


def test_bi_mixture_moments(device):
    """DGFSBiSystem.mixture_moments against the reference writer's direct
    sums (dgfsmomwriterbi.py, 2D) on drifting species states (1e-10)."""
    be, sysm = build(4)
    ov, _, _ = oracles(sysm)
    rng = np.random.default_rng(29)
    fs = []
    for p, nd in enumerate((0.6, 0.4)):
        f = V.species_maxwellian(ov, p, nd, [0.2, -0.1, 0.15], 1.1)[None, :, None] * \
            (1 + 0.1 * rng.random((NUPTS, 1, 6))) * (1 + 0.05 * rng.random((1, ov.n, 1)))
        fs.append(f)
        sysm.species_register(p, 0).set(f)
    vm = sysm.vm
    got = sysm.mixture_moments(0)
    want = _writer_bi_2d(fs, vm.cv(), vm.cw(), vm.masses(), vm.rho0(), vm.u0(), vm.T0(), vm.molarMass0(),
                         vm.R0, vm.NA)
    for p in range(2):
        assert set(got[p]) == set(want[p])
        for k, w in want[p].items():
            err = np.abs(got[p][k] - w).max() / np.abs(w).max()
            assert err < 1e-10, (p, k, err)
    say("bi mixture moments ok", {k: float(got[0][k].mean()) for k in ("rho", "Ux", "T", "Qx", "Pxy")})


def test_bi_annulus_closed_box_conserves_species(device):
    """Triangles of the annular cylinder mesh (straight sides, p = 1) with
    diffuse walls on the cylinder (dgfsbi-wall-diffuse-cyl at rest with a
    temperature varying along it, T = 1 + 0.3 x) and on the far field: every
    particle that hits a wall is re-emitted, and the FP64 projection keeps
    each species' number in Q, so the quadrature-weighted number of each
    species stays constant over two dgfsbi-tvd-rk2 steps (1e-12).  (A wall
    velocity tangential to the circle is not tangential to the polygon's
    sides away from their midpoints; the balance then uses (c - U).n, as the
    reference does, and mass is not conserved exactly.)"""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import bi_couette_cfg
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.mesh import annulus_tris
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    cfg = bi_couette_cfg(Nv=NV, L=LV, Nrho=4, M=12, Mpre=24, masses=MASSES, nden=(0.6, 0.4),
                         omega=0.5, alpha=1.0, prefactor=0.05, order=1, uwall=0.0, project=True)
    cyl = "soln-bcs-cylinder"
    cfg.set(cyl, "type", "dgfsbi-wall-diffuse-cyl")
    cfg.set(cyl, "ux", "0")
    cfg.set(cyl, "uy", "0")
    cfg.set(cyl, "uz", "0")
    cfg.set(cyl, "T", "1 + 0.3*x")
    for far in ("inlet", "outlet"):
        cfg.set(f"soln-bcs-{far}", "type", "dgfsbi-wall-diffuse")
        for k, v in (("T", 1.0), ("ux", 0.0), ("uy", 0.0), ("uz", 0.0)):
            cfg.set(f"soln-bcs-{far}", k, v)
    be = HIPBackend(cfg)
    mesh = annulus_tris(3, 10, r0=0.5, r1=2.0, growth=1.2)
    sysm = DGFSBiSystem(be, mesh, cfg, 6)
    nupts, neles = sysm.ele.nupts, sysm.ele.neles
    djac = sysm.ele.djac_at_np("upts")
    # the p = 1 Williams-Shunn points integrate linears exactly: equal weights
    W = np.full(nupts, 1.0 / nupts) * 2.0

    def number(p):
        f = sysm.species_register(p, 0).get()
        return float(np.einsum("p,pe,pve->", W, djac, f))

    n0 = [number(p) for p in range(2)]
    st = get_stepper("dgfsbi-tvd-rk2", be, sysm, 0.2 * sysm.unit_cfl_dt)
    say("annulus closed box, 2 steps,", neles, "triangles")
    st.advance_to(2 * 0.2 * sysm.unit_cfl_dt)
    be.synchronize()
    for p in range(2):
        n1 = number(p)
        say("  species", p, "number drift", abs(n1 - n0[p]) / n0[p])
        assert abs(n1 - n0[p]) <= 1e-12 * n0[p]
