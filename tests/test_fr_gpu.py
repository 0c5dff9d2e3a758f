"""GPU parity of the FR advection kernels and of whole DGFS steps against the
FP64 oracles (oracle/fr.py for advection, oracle/vhs.py for Q).  FP64 bar:
1e-12 relative for pure advection, 1e-11 for steps that include Q."""
import math
import sys

import numpy as np
import pytest

from oracle import fr as OF
from oracle import vhs as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("m,k,ncol,beta,pad", [(12, 9, 1000, 0.0, 16), (9, 18, 4096, 0.0, 16), (16, 32, 777, 1.0, 16),
                                               (9, 12, 333, 1.0, 16), (6, 12, 64, 0.5, 16), (40, 20, 500, 0.0, 16),
                                               (16, 16, 8256, 0.0, 16), (16, 32, 4128, 1.0, 32),
                                               (12, 9, 1000, 0.0, 15), (16, 64, 640, 0.5, 0),
                                               (9, 6, 130, 1.0, 2)])
def test_mul_mfma(m, k, ncol, beta, pad, device):
    """frfs_mul against A @ B: the paired-column kernel (even ncol and
    leading dimensions, 16-byte aligned operands; full 128-column groups and
    a tail) and the single-column kernel (odd ncol or leading dimension)."""
    import torch
    from dgfs_hip import fr
    say("mul", m, k, ncol, beta, pad)
    rng = np.random.default_rng(m * k)
    A = rng.normal(size=(m, k))
    ldb = ncol + pad
    B = rng.normal(size=(k, ldb))
    C0 = rng.normal(size=(m, ldb))
    tA = torch.as_tensor(A, device=device)
    tB = torch.as_tensor(B, device=device)
    tC = torch.as_tensor(C0.copy(), device=device)
    fr.mul(tA, m, k, k, tB, ldb, tC, ldb, ncol, 1.5, beta, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    C = tC.cpu().numpy()
    want = 1.5 * A @ B[:, :ncol] + beta * C0[:, :ncol]
    assert rel(C[:, :ncol], want) < 1e-14
    assert np.array_equal(C[:, ncol:], C0[:, ncol:])


@pytest.mark.parametrize("offset", [0, 1])
def test_tflux_negdivconf_kernels(offset, device):
    """dgfs_tflux and dgfs_negdivconf on a ragged AoSoA register (37
    elements, 5 velocities, 3 points) against numpy: with 16-byte aligned
    registers (the paired kernels) and with every register one double off
    alignment (the single-element kernels).  Padding elements of the last
    block are left alone by negdivconf."""
    import torch
    from dgfs_hip import fr
    ndims, nupts, neles, nv, vsize, vidx = 2, 3, 37, 5, 9, 3
    nb = (neles + 31) // 32
    ncol = nb * 32 * nv
    ld = ncol + 16
    rng = np.random.default_rng(offset + 1)
    u = rng.normal(size=(nupts, ld))
    sm = rng.normal(size=(ndims * nupts, ld))
    cv = rng.normal(size=(3, vsize))
    rcp = rng.normal(size=(nupts, neles + 1))
    dev = lambda a: torch.as_tensor(np.concatenate([np.zeros(offset), a.ravel()]), device=device)
    tu, tsm, tcv, trcp = dev(u), dev(sm), torch.as_tensor(cv, device=device), dev(rcp)
    tf = dev(np.full((ndims * nupts, ld), 7.0))
    st = torch.cuda.current_stream().cuda_stream
    p = lambda t: t.data_ptr() + 8 * offset
    fr.tflux(ndims, nupts, neles, nv, vidx, p(tu), ld, p(tsm), ld, p(tf), ld, tcv, vsize, st)
    torch.cuda.synchronize()
    f = tf.cpu().numpy()[offset:].reshape(ndims * nupts, ld)
    c = np.arange(ncol)
    blk, lane = c // 32, c % 32
    v, eb = blk % nv, blk // nv
    e = eb * 32 + lane
    want = np.full_like(f, 7.0)
    for i in range(ndims):
        for upt in range(nupts):
            acc = sum(sm[i * nupts + upt][(eb * ndims + k2) * 32 + lane] * cv[k2][vidx + v] for k2 in range(ndims))
            want[i * nupts + upt, :ncol] = acc * u[upt, :ncol]
    say("tflux offset", offset, rel(f, want))
    assert rel(f, want) < 1e-15
    tout = dev(u)
    fr.negdivconf(nupts, neles, nv, p(tout), ld, p(trcp), neles + 1, st)
    torch.cuda.synchronize()
    out = tout.cpu().numpy()[offset:].reshape(nupts, ld)
    want = u.copy()
    live = e < neles
    for upt in range(nupts):
        want[upt, :ncol][live] = -rcp[upt][e[live]] * u[upt, :ncol][live]
    say("negdivconf offset", offset, rel(out, want))
    assert np.array_equal(out, want)


def build_system(order=2, nx=3, ny=2, periodic_x=True, periodic_y=False, Nv=16, L=5.0, NvBS=None,
                 uwall=0.3, nregs=3):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = couette_cfg(Nv=Nv, L=L, Nrho=4, M=6, Kn=1.0, omega=0.5, order=order, uwall=uwall,
                      NvBatchSize=NvBS)
    be = HIPBackend(cfg)
    mesh = rect_quads(nx, ny, 0.0, 0.5 * nx, 0.0, 0.5 * ny, periodic_x=periodic_x, periodic_y=periodic_y)
    return be, DGFSSystem(be, mesh, cfg, nregs), mesh


def test_system_rhs_vs_oracle(device):
    be, sysm, mesh = build_system()
    vm = sysm.vm
    ov = O.VelocityMesh(vm.Nv(), vm.L(), vm.Nrho(), vm.M())
    omesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True)
    walls = {"bottom": (np.array([-0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [-0.3, 0, 0], 1)),
             "top": (np.array([0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [0.3, 0, 0], 1))}
    orc = OF.AdvectionOracle(2, omesh, ov.cv, ov.cw, walls)
    rng = np.random.default_rng(3)
    f = OF.maxwellian(ov.cv, 1.0, [0.1, -0.05, 0], 0.95)[None, :, None] * \
        (1 + 0.1 * rng.random((9, 1, 6))) * (1 + 0.05 * rng.random((1, ov.n, 1)))
    sysm.ele.scal_upts[0].set(f)
    say("system rhs p=2 Nv=16")
    sysm.updatebc(0.0, 0, 1, 0, 0)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    r = sysm.ele.scal_upts[1].get()
    want = orc.rhs(f)
    err = rel(r, want)
    say("  rel L2", err)
    assert err < 1e-12


def test_free_stream_on_gpu(device):
    be, sysm, _ = build_system(periodic_y=True)
    vm = sysm.vm
    m = OF.maxwellian(vm.cv(), 1.0, [0.2, 0.1, 0], 1.0)
    f = np.broadcast_to(m[None, :, None], (9, vm.vsize(), 6)).copy()
    sysm.ele.scal_upts[0].set(f)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    assert np.abs(sysm.ele.scal_upts[1].get()).max() < 1e-12


@pytest.mark.parametrize("stepper,nbs", [("dgfs-euler", None), ("dgfs-tvd-rk2", None), ("dgfs-tvd-rk2", 1024)])
def test_step_vs_oracle(stepper, nbs, device):
    from dgfs_hip.integrators.steppers import get_stepper
    nregs = 2 if stepper == "dgfs-euler" else 3
    be, sysm, mesh = build_system(NvBS=nbs, nregs=nregs)
    vm = sysm.vm
    ov = O.VelocityMesh(vm.Nv(), vm.L(), vm.Nrho(), vm.M())
    omesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True)
    walls = {"bottom": (np.array([-0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [-0.3, 0, 0], 1)),
             "top": (np.array([0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [0.3, 0, 0], 1))}
    orc = OF.AdvectionOracle(2, omesh, ov.cv, ov.cw, walls)
    Qo = O.VHSOracle(ov, 1.0, sysm.sm.prefactor)

    def Q(f):
        fl = f.transpose(2, 0, 1).reshape(-1, ov.n)
        return Qo.q(fl).reshape(6, 9, ov.n).transpose(1, 2, 0)

    rng = np.random.default_rng(4)
    f0 = OF.maxwellian(ov.cv, 1.0, [0.0, 0.0, 0], 1.0)[None, :, None] * (1 + 0.1 * rng.random((9, 1, 6)))
    sysm.eles_scal_upts_full[0].set(f0)
    dt = 0.01
    st = get_stepper(stepper, be, sysm, dt)
    say("step", stepper, "NvBatchSize", vm.NvBatchSize())
    st.advance_to(2 * dt)
    be.synchronize()
    got = sysm.eles_scal_upts_full[st.soln_bank].get()
    f = f0
    for _ in range(2):
        if stepper == "dgfs-euler":
            f = f + dt * (orc.rhs(f) + Q(f))
        else:
            f1 = f + dt * (orc.rhs(f) + Q(f))
            f = 0.5 * f + 0.5 * f1 + 0.5 * dt * (orc.rhs(f1) + Q(f1))
    err = rel(got, f)
    say("  rel L2", err)
    assert err < 1e-11


def test_specular_and_outlet_walls(device):
    """A box with specular walls keeps a state that is even in c_y at rest
    (the reflected ghost equals the state), and an outlet flux equals the
    one-sided upwind flux."""
    import torch
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import base_cfg, maxwellian_section
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = base_cfg(16, 5.0, 4, 6, Kn=1.0, order=2)
    maxwellian_section(cfg, "soln-ics", 1.0, 1.0, (0.2, 0, 0), "maxwellian")
    cfg.set("soln-bcs-bottom", "type", "dgfs-wall-specular")
    cfg.set("soln-bcs-top", "type", "dgfs-wall-specular")
    be = HIPBackend(cfg)
    mesh = rect_quads(3, 2, 0, 1.5, 0, 1, periodic_x=True)
    sysm = DGFSSystem(be, mesh, cfg, 2)
    say("specular walls")
    sysm.ele.scal_upts[0].set(sysm.eles_scal_upts_full[0].get())
    sysm.updatebc(0.0, 0, 1, 0, 0)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    assert np.abs(sysm.ele.scal_upts[1].get()).max() < 1e-12
    del torch


def test_moments_kernel(device):
    """All 15 outputs of the HIP moments kernel against direct numpy sums of
    the reference's definitions (plugins/dgfsmomwriterstd.py:140-214, before
    its dimensional scaling): rho, U, T, q_i = sum f c'_i |c'|^2 cw,
    P_ij = 2 sum f c'_i c'_j cw and p = rho T.  The states are sheared,
    anisotropic and skewed so that every heat-flux and stress component is
    non-zero."""
    be, sysm, _ = build_system()
    vm = sysm.vm
    rng = np.random.default_rng(8)
    cv, cw = vm.cv(), vm.cw()
    base = OF.maxwellian(cv, 1.0, [0.2, -0.1, 0.05], 0.9)
    skew = (1 + 0.3 * np.tanh(cv[0] * cv[1]) + 0.2 * np.tanh(cv[1] * cv[2]) + 0.25 * np.tanh(cv[0] * cv[2])
            + 0.1 * np.tanh(cv[0] ** 3) + 0.1 * np.tanh(cv[1] ** 3) + 0.1 * np.tanh(cv[2] ** 3))
    f = (base * skew)[None, :, None] * (1 + 0.1 * rng.random((9, 1, 6)))
    sysm.eles_scal_upts_full[0].set(f)
    mom = sysm.moments(0)
    rho = f.sum(1) * cw
    U = np.stack([(f * cv[d][None, :, None]).sum(1) * cw / rho for d in range(3)])
    c = cv[:, None, :, None] - U[:, :, None, :]
    c2 = (c ** 2).sum(0)
    T = (f * c2).sum(1) * cw * 2 / 3 / rho
    assert np.allclose(mom[:, 0], rho, rtol=1e-13)
    assert np.allclose(mom[:, 1:4], U.transpose(1, 0, 2), rtol=1e-12, atol=1e-14)
    assert np.allclose(mom[:, 4], T, rtol=1e-12)
    for i in range(3):
        qi = (f * c[i] * c2).sum(1) * cw
        assert np.abs(qi).min() > 1e-4
        assert np.allclose(mom[:, 5 + i], qi, rtol=1e-10, atol=1e-13), i
    for m, (i, j) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        pij = 2 * (f * c[i] * c[j]).sum(1) * cw
        assert np.abs(pij).min() > 1e-4
        assert np.allclose(mom[:, 8 + m], pij, rtol=1e-11, atol=1e-13), (i, j)
    # the trace of P is 3 rho T (T = (2/3) <|c'|^2>)
    assert np.allclose(mom[:, 8] + mom[:, 11] + mom[:, 13], 3 * rho * T, rtol=1e-12)
    assert np.allclose(mom[:, 14], rho * T, rtol=1e-12)


def test_triangle_mesh_equilibrium_and_mass(device):
    """Triangle elements (p = 2, Williams-Shunn points): a wall-equilibrium
    state is a steady state of the advection (rhs = 0), and the advection of
    a perturbed state conserves the total mass (x-periodic, diffuse walls)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg
    from dgfs_hip.mesh import rect_tris
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=2, uwall=0.0)
    cfg.set("solver-elements-tri", "soln-pts", "williams-shunn")
    be = HIPBackend(cfg)
    mesh = rect_tris(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    sysm = DGFSSystem(be, mesh, cfg, 3)
    vm = sysm.vm
    m = OF.maxwellian(vm.cv(), 1.0, [0.0, 0.0, 0.0], 1.0)
    nupts, neles = sysm.ele.nupts, sysm.ele.neles
    f = np.broadcast_to(m[None, :, None], (nupts, vm.vsize(), neles)).copy()
    sysm.ele.scal_upts[0].set(f)
    sysm.updatebc(0.0, 0, 1, 0, 0)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    r = sysm.ele.scal_upts[1].get()
    say("tri equilibrium max |rhs|", np.abs(r).max())
    assert np.abs(r).max() < 1e-11 * np.abs(f).max() * 10
    # perturbed state: the advection conserves the total mass (sum over
    # velocities of the element integrals with the nodal quadrature W_p |J|):
    # interior fluxes cancel pairwise and the diffuse walls reflect every
    # incoming particle
    rng = np.random.default_rng(8)
    f1 = f * (1 + 0.05 * rng.random((nupts, 1, neles))) * (1 + 0.05 * rng.random((1, vm.vsize(), 1)))
    sysm.ele.scal_upts[0].set(f1)
    sysm.updatebc(0.0, 0, 1, 0, 0)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    r = sysm.ele.scal_upts[1].get()
    djac = sysm.ele.djac_at_np("upts")          # (nupts, neles)
    W = _tri_nodal_weights(sysm.ele)            # (nupts,)
    rate = float(np.einsum("p,pe,pve->", W, djac, r))
    scale = float(np.einsum("p,pe,pve->", W, djac, np.abs(r)))
    say("tri mass rate / int |rhs|", abs(rate) / scale)
    assert abs(rate) < 1e-12 * scale


def _tri_nodal_weights(ele):
    """W_p = integral over the reference triangle of the nodal basis function
    l_p (collapsed Gauss-Legendre rule of 12 x 12 points, exact for the
    degree-2 basis)."""
    a, wa = np.polynomial.legendre.leggauss(12)
    A, B = np.meshgrid(a, a, indexing="ij")
    WA = np.outer(wa, wa) * (1 - B) / 2
    xi = (1 + A) * (1 - B) / 2 - 1
    pts = np.stack([xi.ravel(), B.ravel()], axis=1)
    L = ele.basis.ubasis.at(pts)                 # (npts, nupts)
    return WA.ravel() @ L


def test_residual_plugin(device):
    """DGFSResidualStdPlugin (plugins/dgfsresidualstd.py:9-80): every nsteps
    steps sqrt(sum (f_n - f_{n-1})^2 / sum f_{n-1}^2) over the valid entries
    of the register, from the HIP reduction, against numpy on the unpacked
    registers (1e-12)."""
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.plugins import DGFSResidualStdPlugin
    be, sysm, _ = build_system(nregs=3)
    rng = np.random.default_rng(9)
    f0 = sysm.eles_scal_upts_full[0].get()
    sysm.eles_scal_upts_full[0].set(f0 * (1 + 0.05 * rng.random(f0.shape)))
    st = get_stepper("dgfs-tvd-rk2", be, sysm, 0.01)
    seen = []
    st.plugins.append(lambda s: seen.append(s.system.eles_scal_upts_full[s.soln_bank].get()))
    pl = DGFSResidualStdPlugin(st, 2)
    st.plugins.append(pl)
    say("residual plugin, 4 steps")
    st.advance_to(0.04)
    assert len(pl.rows) == 2
    for k, (i, j) in enumerate([(0, 1), (2, 3)]):
        want = math.sqrt(np.sum((seen[j] - seen[i]) ** 2) / np.sum(seen[i] ** 2))
        say("  ", pl.rows[k][1], want)
        assert abs(pl.rows[k][1] - want) <= 1e-12 * want


def _channel_oracle(vm, bottom_wall=None):
    ov = O.VelocityMesh(vm.Nv(), vm.L(), vm.Nrho(), vm.M())
    omesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True)
    walls = {"bottom": (np.array([-0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [-0.3, 0, 0], 1)),
             "top": (np.array([0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [0.3, 0, 0], 1))}
    if bottom_wall is not None:
        walls["bottom"] = bottom_wall
    return ov, OF.AdvectionOracle(2, omesh, ov.cv, ov.cw, walls)


def _perturbed(ov, seed):
    rng = np.random.default_rng(seed)
    return OF.maxwellian(ov.cv, 1.0, [0.1, -0.05, 0], 0.95)[None, :, None] * \
        (1 + 0.1 * rng.random((9, 1, 6))) * (1 + 0.05 * rng.random((1, ov.n, 1)))


def _rhs_of(be, sysm, f):
    sysm.ele.scal_upts[0].set(f)
    sysm.updatebc(0.0, 0, 1, 0, 0)
    sysm.rhs(0.0, 0, 1, 0, 0)
    be.synchronize()
    return sysm.ele.scal_upts[1].get()


def test_cyl_wall_vs_oracle(device):
    """dgfs-wall-diffuse-cyl (inters.py:446-492): a wall whose velocity and
    temperature are expressions of x, against the oracle's per-point diffuse
    wall (1e-12)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import U0, couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=2, uwall=0.3)
    cfg.set("constants", "uw", 0.3 * U0)
    sect = "soln-bcs-bottom"
    cfg.set(sect, "type", "dgfs-wall-diffuse-cyl")
    cfg.set(sect, "ux", "uw*sin(2*pi*x/1.5)")
    cfg.set(sect, "uy", "0")
    cfg.set(sect, "uz", "0")
    cfg.set(sect, "T", "1 + 0.2*x")
    be = HIPBackend(cfg)
    sysm = DGFSSystem(be, rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True), cfg, 3)

    def wall(xy):
        U = np.zeros((len(xy), 3))
        U[:, 0] = 0.3 * np.sin(2 * np.pi * xy[:, 0] / 1.5)
        return U, 1.0 + 0.2 * xy[:, 0]

    ov, orc = _channel_oracle(sysm.vm, wall)
    f = _perturbed(ov, 11)
    say("cyl wall p=2 Nv=16")
    r = _rhs_of(be, sysm, f)
    err = rel(r, orc.rhs(f))
    say("  rel L2", err)
    assert err < 1e-12


def test_gmsh_mesh_system_vs_oracle(device, tmp_path):
    """A channel read from a gmsh file (periodic x pair, named walls) runs
    the same advection as the oracle on the structured mesh (1e-12)."""
    from test_gmsh import write_msh
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg
    from dgfs_hip.gmsh import read_gmsh
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    walls = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0)
    path = tmp_path / "channel.msh"
    path.write_text("\n".join(write_msh(walls, ["bottom", "top", "left", "right"],
                                        rename={"left": "periodic_0_r", "right": "periodic_0_l"},
                                        shuffle_seed=1)) + "\n")
    mesh = read_gmsh(str(path))
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=2, uwall=0.3)
    be = HIPBackend(cfg)
    sysm = DGFSSystem(be, mesh, cfg, 3)
    ov, orc = _channel_oracle(sysm.vm)
    f = _perturbed(ov, 12)
    say("gmsh channel p=2 Nv=16")
    r = _rhs_of(be, sysm, f)
    err = rel(r, orc.rhs(f))
    say("  rel L2", err)
    assert err < 1e-12


def test_gmsh_triangles_match_structured(device, tmp_path):
    """The triangle channel read back from a gmsh file (periodic pair, named
    walls) gives the same advection right-hand side as the structured
    triangle mesh it was written from (bitwise: same elements, faces and
    pairings)."""
    from test_gmsh import write_msh
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg
    from dgfs_hip.gmsh import read_gmsh
    from dgfs_hip.mesh import rect_tris
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    walls = rect_tris(3, 2, 0.0, 1.5, 0.0, 1.0)
    path = tmp_path / "tris.msh"
    path.write_text("\n".join(write_msh(walls, ["bottom", "top", "left", "right"],
                                        rename={"left": "periodic_0_r", "right": "periodic_0_l"})) + "\n")
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=2, uwall=0.3)
    out = []
    for mesh in (rect_tris(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True), read_gmsh(str(path))):
        be = HIPBackend(cfg)
        sysm = DGFSSystem(be, mesh, cfg, 3)
        rng = np.random.default_rng(5)
        f = sysm.eles_scal_upts_full[0].get() * (1 + 0.1 * rng.random((sysm.ele.nupts, 1, sysm.ele.neles)))
        out.append(_rhs_of(be, sysm, f))
    say("gmsh triangles vs structured: max |diff|", float(np.abs(out[0] - out[1]).max()))
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("order", [1, 2])
def test_triangle_advection_converges(order, device):
    """Triangle FR (Williams-Shunn points) on the GPU, checked by the
    mathematics: a state f = 1 + 0.1 sin(2 pi x), uniform in y and even in
    c_y, between flat specular walls (whose ghost state then equals the
    state) in an x-periodic channel, has the exact right-hand side
    -c_x 0.2 pi cos(2 pi x); the pointwise error of this derivative falls
    at order p under refinement (measured rate above p - 0.3, the bar of the
    quadrilateral oracle's convergence pin)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import base_cfg
    from dgfs_hip.mesh import rect_tris
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    errs = []
    for nx in (4, 8, 16):
        cfg = base_cfg(4, 3.0, 2, 6, Kn=1.0, order=order)
        cfg.set("soln-bcs-bottom", "type", "dgfs-wall-specular")
        cfg.set("soln-bcs-top", "type", "dgfs-wall-specular")
        be = HIPBackend(cfg)
        nupts, neles = (order + 1) * (order + 2) // 2, 2 * nx * 2
        sysm = DGFSSystem(be, rect_tris(nx, 2, 0.0, 1.0, 0.0, 0.5, periodic_x=True), cfg, 2,
                          initsoln=np.ones((nupts, 64, neles)))
        xy = np.asarray(sysm.ele.ploc_at_np("upts"))          # (nupts, neles, 2)
        cv = sysm.vm.cv()
        x = xy[:, :, 0]
        f = (1 + 0.1 * np.sin(2 * np.pi * x))[:, None, :] * np.ones(cv.shape[1])[None, :, None]
        exact = -cv[0][None, :, None] * (0.2 * np.pi * np.cos(2 * np.pi * x))[:, None, :]
        r = _rhs_of(be, sysm, f)
        errs.append(float(np.abs(r - exact).max()))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    say("triangle advection p =", order, "errors", errs, "rates", rates)
    assert np.all(rates > order - 0.3), (errs, rates)


@pytest.mark.parametrize("order", [1, 3])
def test_system_rhs_vs_oracle_orders(order, device):
    """The RHS at p = 1 and at p = 3 (the cavity's order: 16 solution points,
    M1 - M3 M2 of 16 x 32) against the quadrilateral oracle (1e-12)."""
    be, sysm, _ = build_system(order=order)
    vm = sysm.vm
    ov = O.VelocityMesh(vm.Nv(), vm.L(), vm.Nrho(), vm.M())
    omesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True)
    walls = {"bottom": (np.array([-0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [-0.3, 0, 0], 1)),
             "top": (np.array([0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [0.3, 0, 0], 1))}
    orc = OF.AdvectionOracle(order, omesh, ov.cv, ov.cw, walls)
    nupts = (order + 1) ** 2
    rng = np.random.default_rng(30 + order)
    f = OF.maxwellian(ov.cv, 1.0, [0.1, -0.05, 0], 0.95)[None, :, None] * \
        (1 + 0.1 * rng.random((nupts, 1, 6))) * (1 + 0.05 * rng.random((1, ov.n, 1)))
    say("system rhs p =", order)
    r = _rhs_of(be, sysm, f)
    err = rel(r, orc.rhs(f))
    say("  rel L2", err)
    assert err < 1e-12


@pytest.mark.parametrize("order,walls", [(1, "diffuse"), (2, "diffuse"), (2, "farfield")])
def test_triangle_rhs_vs_oracle(order, walls, device):
    """Triangles (Williams-Shunn points) on the x-periodic channel, against the
    triangle oracle built from monomials and the definition of the DG lift
    (oracle/fr.py TriFR): diffuse moving walls, or an inlet at the bottom and
    an outlet at the top (1e-12)."""
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg, maxwellian_section
    from dgfs_hip.mesh import rect_tris
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=order, uwall=0.3)
    ov = O.VelocityMesh(16, 5.0, 4, 6)
    if walls == "farfield":
        maxwellian_section(cfg, "soln-bcs-bottom", 1.2, 0.9, (0.1, 0.6, 0.0), "dgfs-inlet")
        cfg._cp.remove_section("soln-bcs-top")
        cfg.set("soln-bcs-top", "type", "dgfs-outlet")
        owalls = {"bottom": ("inlet", np.array([0.1, 0.6, 0.0]), OF.maxwellian(ov.cv, 1.2, [0.1, 0.6, 0.0], 0.9)),
                  "top": ("outlet",)}
    else:
        owalls = {"bottom": (np.array([-0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [-0.3, 0, 0], 1)),
                  "top": (np.array([0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [0.3, 0, 0], 1))}
    be = HIPBackend(cfg)
    mesh = rect_tris(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    sysm = DGFSSystem(be, mesh, cfg, 2)
    verts = np.transpose(mesh.nodes, (1, 0, 2))
    orc = OF.TriMeshAdvectionOracle(order, verts, ov.cv, ov.cw, walls=owalls, period=1.5,
                                    boundary_name=lambda m: "bottom" if abs(m[1]) < 1e-9 else "top")
    nupts, neles = sysm.ele.nupts, sysm.ele.neles
    rng = np.random.default_rng(40 + order)
    f = OF.maxwellian(ov.cv, 1.0, [0.1, -0.05, 0], 0.95)[None, :, None] * \
        (1 + 0.1 * rng.random((nupts, 1, neles))) * (1 + 0.2 * np.tanh(ov.cv[0] - ov.cv[1]))[None, :, None]
    say("triangle rhs p =", order, walls)
    r = _rhs_of(be, sysm, f)
    err = rel(r, orc.rhs(f))
    say("  rel L2", err)
    assert err < 1e-12


def test_rhs_through_the_provider_walk(device):
    """The advection RHS assembled only from backend.kernel(...) calls with
    the reference's kernel names and argument names (elements.py:56-164,
    inters.py:17-58 and 232-309: 'mul', 'tflux', 'intcflux', 'bccupdate',
    'bccflux', 'negdivconf'), run through a backend queue, equals the
    oracle (1e-12) and the system's own rhs (bitwise)."""
    be, sysm, _ = build_system()
    ele, vm, iint = sysm.ele, sysm.vm, sysm.iint
    for mod in ("tflux", "negdivconf", "intcflux", "bccflux", "bccupdate"):
        be.pointwise.register(f"frfs.solvers.dgfs.kernels.{mod}")
    cvs = dict(cvx=vm.d_cvx(), cvy=vm.d_cvy(), cvz=vm.d_cvz())
    nupts, neles, nvars = ele.nupts, ele.neles, ele.nvars
    disu = be.kernel("mul", ele.opmat("M0"), ele.scal_upts_inb, out=ele._scal_fpts)
    tdisf = be.kernel("tflux", tplargs=dict(ndims=2, nvars=nvars, vsize=vm.vsize()), dims=[nupts, neles],
                      u=ele.scal_upts_inb, smats=ele._smats_upts, f=ele._vect_upts, **cvs)
    tdivtpcorf = be.kernel("mul", ele.opmat("M1 - M3*M2"), ele._vect_upts, out=ele.scal_upts_outb)
    tdivtconf = be.kernel("mul", ele.opmat("M3"), ele._scal_fpts, out=ele.scal_upts_outb, beta=1.0)
    negdiv = be.kernel("negdivconf", tplargs=dict(ndims=2, nvars=nvars, srcex=[]), dims=[nupts, neles],
                       tdivtconf=ele.scal_upts_outb, rcpdjac=ele._rcpdjac_upts)
    icomm = be.kernel("intcflux", tplargs=dict(ndims=2, nvars=nvars, rsolver="rusanov", vsize=vm.vsize()),
                      dims=[iint.ninterfpts], ul=iint._scal_lhs, ur=iint._scal_rhs, magnl=iint._magnl,
                      nl=iint._nl, **cvs)
    bcomm, bupd = [], []
    for b in sysm.bcints:
        tpl = dict(ndims=2, nvars=nvars, rsolver="rusanov", bctype=b.type, vsize=vm.vsize(), cw=vm.cw(),
                   c=dict(Ux=b._U[0], Uy=b._U[1], Uz=b._U[2]))
        bcomm.append(be.kernel("bccflux", tplargs=tpl, dims=[b.ninterfpts], ul=b._scal_lhs, magnl=b._magnl,
                               nl=b._nl, bnd_f0=b._f0, bc_vals=b._bcvals, **cvs))
        bupd.append(be.kernel("bccupdate", tplargs=tpl, dims=[b.ninterfpts], ul=b._scal_lhs, nl=b._nl,
                              bnd_f0=b._f0, bc_vals=b._bcvals, **cvs))
    ov = O.VelocityMesh(vm.Nv(), vm.L(), vm.Nrho(), vm.M())
    orc = OF.AdvectionOracle(2, OF.RectMesh(3, 2, 0, 1.5, 0, 1.0, periodic_x=True), ov.cv, ov.cw, {
        "bottom": (np.array([-0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [-0.3, 0, 0], 1)),
        "top": (np.array([0.3, 0, 0]), OF.maxwellian(ov.cv, 1, [0.3, 0, 0], 1))})
    f = _perturbed(ov, 21)
    ele.scal_upts[0].set(f)
    ele.scal_upts_inb.active, ele.scal_upts_outb.active = 0, 1
    q = be.queue()
    q << disu()
    q << [k(t=0.0, varidx=0) for k in bupd]
    be.runall([q])
    q << disu()
    q << tdisf(varidx=0)
    q << tdivtpcorf()
    q << icomm(varidx=0)
    q << [k(t=0.0, varidx=0) for k in bcomm]
    q << tdivtconf()
    q << negdiv(t=0.0)
    be.runall([q])
    be.synchronize()
    walk = ele.scal_upts[1].get()
    err = rel(walk, orc.rhs(f))
    say("rhs through the provider walk: rel L2 vs oracle", err)
    assert err < 1e-12
    assert np.array_equal(walk, _rhs_of(be, sysm, f))
