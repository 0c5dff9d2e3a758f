"""GPU parity of the far-field and specular boundaries against the FR oracle's
ghost states (oracle/fr.py, pinned by the inflow/outflow balance in
tests/test_oracle_fr.py): dgfs-inlet, dgfs-inlet-normalshock, dgfs-outlet,
dgfs-vanish and dgfs-wall-specular (frfs/solvers/dgfs/inters.py:232-443,
kernels/bcs/dgfs-*.mako), and their dgfsbi- forms with each species'
inflow Maxwellian (frfs/solvers/dgfsbi/inters.py:424-676).

The channel is periodic in y and bounded in x, so the x = 0 and x = 1.5
faces are the boundaries; the states are perturbed in space and velocity so
that every ghost branch carries weight.  Bar: 1e-12 relative L2 of the
advection right-hand side (FP64)."""
import sys

import numpy as np
import pytest

from oracle import fr as OF
from oracle import vhs as O
from oracle import vss as V

pytestmark = pytest.mark.gpu

# the inflow Maxwellian's velocity: at Nv = 16 on [-5, 5]^3 the grid value
# c_x = 0.3125 lies between 0 and U_x, where the inlet and normal-shock
# boundaries differ
UIN = (0.7, 0.05, 0.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


PAIRS = [("dgfs-inlet", "dgfs-outlet"), ("dgfs-inlet-normalshock", "dgfs-vanish"),
         ("dgfs-wall-specular", "dgfs-outlet"), ("dgfs-vanish", "dgfs-inlet-normalshock"),
         ("dgfs-outlet", "dgfs-wall-specular")]


def _section(cfg, sect, kind, rho, T):
    from dgfs_hip.cases import maxwellian_section
    if kind in ("dgfs-inlet", "dgfs-inlet-normalshock"):
        maxwellian_section(cfg, sect, rho, T, UIN, kind)
    else:
        cfg.set(sect, "type", kind)


def _oracle_wall(kind, cv, rho, T):
    f0 = OF.maxwellian(cv, rho, UIN, T)
    return {"dgfs-inlet": ("inlet", np.array(UIN), f0), "dgfs-inlet-normalshock": ("inlet-normalshock", f0),
            "dgfs-outlet": ("outlet",), "dgfs-vanish": ("vanish",), "dgfs-wall-specular": ("specular",)}[kind]


@pytest.mark.parametrize("left,right", PAIRS)
def test_farfield_rhs_vs_oracle(left, right, device):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import base_cfg, maxwellian_section
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    Nv, L = 16, 5.0
    cfg = base_cfg(Nv, L, 4, 6, Kn=1.0, order=2)
    maxwellian_section(cfg, "soln-ics", 1.0, 1.0, (0.2, 0, 0), "maxwellian")
    _section(cfg, "soln-bcs-left", left, 1.3, 0.8)
    _section(cfg, "soln-bcs-right", right, 0.9, 1.2)
    be = HIPBackend(cfg)
    mesh = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_y=True)
    sysm = DGFSSystem(be, mesh, cfg, 2)
    ov = O.VelocityMesh(Nv, L, 4, 6)
    assert np.any((ov.cv[0] > 0) & (ov.cv[0] < UIN[0]))
    orc = OF.AdvectionOracle(2, OF.RectMesh(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=False, periodic_y=True),
                             ov.cv, ov.cw, walls={"left": _oracle_wall(left, ov.cv, 1.3, 0.8),
                                                  "right": _oracle_wall(right, ov.cv, 0.9, 1.2)})
    rng = np.random.default_rng(PAIRS.index((left, right)))
    f = OF.maxwellian(ov.cv, 1.0, [0.25, -0.1, 0.05], 0.95)[None, :, None] * \
        (1 + 0.1 * rng.random((9, 1, 6))) * (1 + 0.3 * np.tanh(ov.cv[0] + 0.5 * ov.cv[1]))[None, :, None]
    say("far field", left, right)
    sysm.ele.scal_upts[0].set(f)
    sysm.updatebc(0.0, 0, 1, 0, 0)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    r = sysm.ele.scal_upts[1].get()
    err = rel(r, orc.rhs(f))
    say("  rel L2", err)
    assert err < 1e-12


BI_MASSES = (1.0, 2.1)
BI_NDEN = (0.6, 0.4)


@pytest.mark.parametrize("left,right", [("dgfsbi-inlet", "dgfsbi-outlet"),
                                        ("dgfsbi-inlet-normalshock", "dgfsbi-vanish"),
                                        ("dgfsbi-wall-specular", "dgfsbi-inlet")])
def test_bi_farfield_rhs_vs_oracle(left, right, device):
    """Two species at 16^3 on [-4, 4]^3 (the heavier species' inflow
    Maxwellian passes the 1e-5 moment check there): every species' ghost is
    its own inflow Maxwellian n_p (m_p/(pi T))^(3/2) exp(-m_p |c - U|^2/T)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import _bi_state, bi_couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    Nv, L = 16, 4.0
    cfg = bi_couette_cfg(Nv=Nv, L=L, Nrho=4, M=12, Mpre=24, masses=BI_MASSES, nden=BI_NDEN, omega=0.5,
                         alpha=1.0, prefactor=0.05, order=1)
    for sect in ("soln-bcs-bottom", "soln-bcs-top"):
        cfg._cp.remove_section(sect)
    states = {"soln-bcs-left": (1.1, (0.5, 0.05, 0.0), (0.7, 0.5)),
              "soln-bcs-right": (0.9, (-0.3, 0.0, 0.1), (0.4, 0.8))}
    for sect, kind in (("soln-bcs-left", left), ("soln-bcs-right", right)):
        T, u, nd = states[sect]
        if kind in ("dgfsbi-inlet", "dgfsbi-inlet-normalshock"):
            _bi_state(cfg, sect, T, u, nd, kind=kind)
        else:
            cfg.set(sect, "type", kind)
    be = HIPBackend(cfg)
    sysm = DGFSBiSystem(be, rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_y=True), cfg, 4)
    ov = V.BiVelocityMesh(Nv, L, 4, 12, BI_MASSES)
    omesh = OF.RectMesh(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=False, periodic_y=True)
    rng = np.random.default_rng(17)
    for p in range(2):
        walls = {}
        for sect, name, kind in (("soln-bcs-left", "left", left), ("soln-bcs-right", "right", right)):
            T, u, nd = states[sect]
            f0 = V.species_maxwellian(ov, p, nd[p], list(u), T)
            walls[name] = {"dgfsbi-inlet": ("inlet", np.array(u), f0),
                           "dgfsbi-inlet-normalshock": ("inlet-normalshock", f0),
                           "dgfsbi-outlet": ("outlet",), "dgfsbi-vanish": ("vanish",),
                           "dgfsbi-wall-specular": ("specular",)}[kind]
        orc = OF.AdvectionOracle(1, omesh, ov.cv, ov.cw, walls)
        f = V.species_maxwellian(ov, p, BI_NDEN[p], [0.1, -0.05, 0.0], 1.0)[None, :, None] * \
            (1 + 0.1 * rng.random((4, 1, 6))) * (1 + 0.3 * np.tanh(ov.cv[0] - 0.4 * ov.cv[1]))[None, :, None]
        sysm.ele.scal_upts[0].set(f)
        sysm.updatebc(0.0, p, 0, 1, 0, 0)
        sysm.rhs(0.0, p, 0, 1, 0, 0)
        be.synchronize()
        err = rel(sysm.ele.scal_upts[1].get(), orc.rhs(f))
        say("bi far field", left, right, "species", p, "rel L2", err)
        assert err < 1e-12
