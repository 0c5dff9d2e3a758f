"""Pins of the FR oracle (oracle/fr.py) and of the product's host-side FR
operators against it.  CPU only."""
import numpy as np
import pytest

from oracle import fr as OF
from oracle import vhs as O


@pytest.mark.parametrize("p", [1, 2, 3])
def test_interpolation_and_derivative_exact(p):
    q = OF.QuadFR(p)
    rng = np.random.default_rng(p)
    c = rng.normal(size=(p + 1, p + 1))
    poly = lambda x, y: sum(c[i, j] * x ** i * y ** j for i in range(p + 1) for j in range(p + 1))
    dpx = lambda x, y: sum(c[i, j] * i * x ** max(i - 1, 0) * y ** j for i in range(1, p + 1) for j in range(p + 1))
    u = poly(q.upts[:, 0], q.upts[:, 1])
    assert np.allclose(q.basis_at(q.fpts) @ u, poly(q.fpts[:, 0], q.fpts[:, 1]), atol=1e-12)
    assert np.allclose(q.grad_at(q.upts)[0] @ u, dpx(q.upts[:, 0], q.upts[:, 1]), atol=1e-11)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_lift_identity(p):
    q = OF.QuadFR(p)
    M3 = q.lift()
    zq, wq = OF.gauss_lobatto(p + 1)
    rng = np.random.default_rng(7)
    c = rng.normal(size=(p + 1, p + 1))
    poly = lambda x, y: sum(c[i, j] * x ** i * y ** j for i in range(p + 1) for j in range(p + 1))
    lk = OF.lagrange(q.z, zq)
    for f, (proj, _) in enumerate(q.faces):
        xs, ys = proj(zq)
        for k in range(p + 1):
            lhs = np.sum(q.W * M3[:, f * (p + 1) + k] * poly(q.upts[:, 0], q.upts[:, 1]))
            rhs = np.sum(wq * lk[:, k] * poly(xs, ys))
            assert abs(lhs - rhs) < 1e-12 * max(1, abs(rhs))


def test_gauss_lobatto_table():
    z, w = OF.gauss_lobatto(4)
    assert np.allclose(z, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1]) and np.allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6])


@pytest.mark.parametrize("p", [2, 3])
def test_product_operators_match_oracle(p):
    from dgfs_hip.shapes import get_shape
    s = get_shape("quad", p)
    q = OF.QuadFR(p)
    assert np.allclose(s.upts, q.upts, atol=1e-15) and np.allclose(s.fpts, q.fpts, atol=1e-15)
    assert np.allclose(s.m0, q.basis_at(q.fpts), atol=1e-13)
    D = q.grad_at(q.upts)
    assert np.allclose(s.m1, np.concatenate([D[0], D[1]], axis=1), atol=1e-12)
    assert np.allclose(s.m3, q.lift(), atol=1e-12)
    M2 = np.concatenate([q.fnorm[:, 0:1] * q.basis_at(q.fpts), q.fnorm[:, 1:2] * q.basis_at(q.fpts)], axis=1)
    assert np.allclose(s.opmat("M1 - M3*M2"), np.concatenate([D[0], D[1]], axis=1) - q.lift() @ M2, atol=1e-12)


def test_tri_operators_basic():
    from dgfs_hip.shapes import get_shape
    s = get_shape("tri", 2)
    x, y = s.upts[:, 0], s.upts[:, 1]
    u = 1 + 2 * x - y + 0.5 * x * y + 0.3 * y * y
    fx, fy = s.fpts[:, 0], s.fpts[:, 1]
    assert np.allclose(s.m0 @ u, 1 + 2 * fx - fy + 0.5 * fx * fy + 0.3 * fy * fy, atol=1e-12)
    du = s.m1 @ np.concatenate([u, np.zeros_like(u)])
    assert np.allclose(du, 2 + 0.5 * y, atol=1e-11)


def _vm(N=8, L=5.0):
    return O.VelocityMesh(N, L, 2, 3)


def test_free_stream_preserved():
    vm = _vm()
    mesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1, periodic_x=True, periodic_y=True)
    orc = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw)
    f = np.broadcast_to(OF.maxwellian(vm.cv, 1.0, [0.2, 0.1, 0], 1.0)[None, :, None], (9, vm.n, 6)).copy()
    assert np.abs(orc.rhs(f)).max() < 1e-12


def test_periodic_conservation_per_velocity():
    vm = _vm()
    mesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1, periodic_x=True, periodic_y=True)
    orc = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw)
    rng = np.random.default_rng(1)
    f = rng.random((9, vm.n, 6))
    r = orc.rhs(f)
    tot = np.einsum("p,pve->v", orc.fr.W, r) * orc.djac
    assert np.abs(tot).max() < 1e-11 * np.abs(r).max()


def test_diffuse_wall_zero_net_mass_flux():
    vm = _vm(12, 5.0)
    mesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1, periodic_x=True)
    fw0 = OF.maxwellian(vm.cv, 1.0, [0.3, 0, 0], 1.0)
    fw1 = OF.maxwellian(vm.cv, 1.0, [-0.3, 0, 0], 1.0)
    orc = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw,
                             walls={"bottom": (np.array([0.3, 0, 0]), fw0), "top": (np.array([-0.3, 0, 0]), fw1)})
    rng = np.random.default_rng(2)
    f = OF.maxwellian(vm.cv, 1.0, [0, 0, 0], 0.9)[None, :, None] * (1 + 0.2 * rng.random((9, 1, 6)))
    r = orc.rhs(f)
    mass = np.einsum("p,pve->", orc.fr.W, r) * orc.djac * vm.cw
    assert abs(mass) < 1e-12 * np.abs(r).sum() * vm.cw


def _cyl_state(vm, seed):
    rng = np.random.default_rng(seed)
    return OF.maxwellian(vm.cv, 1.0, [0.1, 0, 0], 0.9)[None, :, None] * (1 + 0.2 * rng.random((9, 1, 6)))


def test_cyl_wall_constant_equals_diffuse():
    """The unnormalised wall Maxwellian of the varying wall with constant U, T
    gives the diffuse wall exactly: the normalisation cancels in the balance."""
    vm = _vm(12, 5.0)
    mesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1, periodic_x=True)
    U0, U1 = np.array([0.3, 0, 0]), np.array([-0.2, 0, 0])
    top = (U1, OF.maxwellian(vm.cv, 1.0, U1, 1.3))
    a = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw,
                           walls={"bottom": (U0, OF.maxwellian(vm.cv, 1.0, U0, 0.8)), "top": top})
    b = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw, walls={
        "bottom": lambda xy: (np.tile(U0, (len(xy), 1)), np.full(len(xy), 0.8)), "top": top})
    f = _cyl_state(vm, 5)
    ra, rb = a.rhs(f), b.rhs(f)
    assert np.abs(ra - rb).max() < 1e-13 * np.abs(ra).max()


def test_cyl_wall_uses_the_point_location():
    """A wall velocity that is piecewise constant in x gives, element by
    element, the rhs of the constant wall that holds there."""
    vm = _vm(12, 5.0)
    mesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1, periodic_x=True)
    top = (np.zeros(3), OF.maxwellian(vm.cv, 1.0, [0, 0, 0], 1.0))
    Ua, Ub = np.array([0.3, 0, 0]), np.array([-0.3, 0, 0])

    def wall(xy):
        U = np.where((xy[:, 0] < 0.5)[:, None], Ua, Ub)
        return U, np.where(xy[:, 0] < 0.5, 1.0, 1.2)

    f = _cyl_state(vm, 6)
    r = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw, walls={"bottom": wall, "top": top}).rhs(f)
    ra = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw,
                            walls={"bottom": (Ua, OF.maxwellian(vm.cv, 1, Ua, 1.0)), "top": top}).rhs(f)
    rb = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw,
                            walls={"bottom": (Ub, OF.maxwellian(vm.cv, 1, Ub, 1.2)), "top": top}).rhs(f)
    # bottom row: element 0 lies at x < 0.5, elements 1 and 2 beyond
    assert np.abs(r[..., 0] - ra[..., 0]).max() < 1e-13 * np.abs(ra).max()
    assert np.abs(r[..., 1:3] - rb[..., 1:3]).max() < 1e-13 * np.abs(rb).max()
    assert np.abs(r[..., 0] - rb[..., 0]).max() > 1e-3 * np.abs(rb).max()


def test_cyl_wall_zero_net_mass_flux():
    vm = _vm(12, 5.0)
    mesh = OF.RectMesh(3, 2, 0, 1.5, 0, 1, periodic_x=True)

    def wall(xy):
        U = np.zeros((len(xy), 3))
        U[:, 0] = 0.3 * np.sin(2 * np.pi * xy[:, 0] / 1.5)
        return U, 1.0 + 0.2 * xy[:, 0]

    orc = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw, walls={"bottom": wall, "top": wall})
    r = orc.rhs(_cyl_state(vm, 7))
    mass = np.einsum("p,pve->", orc.fr.W, r) * orc.djac * vm.cw
    assert abs(mass) < 1e-12 * np.abs(r).sum() * vm.cw


def test_wall_equilibrium_rest():
    vm = _vm(12, 5.0)
    mesh = OF.RectMesh(2, 2, 0, 1, 0, 1, periodic_x=True)
    fw = OF.maxwellian(vm.cv, 1.0, [0, 0, 0], 1.0)
    orc = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw,
                             walls={"bottom": (np.zeros(3), fw), "top": (np.zeros(3), fw)})
    f = np.broadcast_to(fw[None, :, None], (9, vm.n, 4)).copy()
    assert np.abs(orc.rhs(f)).max() < 1e-12 * fw.max()


@pytest.mark.parametrize("p", [2, 3])
def test_advection_convergence(p):
    vm = _vm(4, 3.0)
    errs = []
    for nx in (4, 8, 16):
        mesh = OF.RectMesh(nx, 1, 0, 1, 0, 1, periodic_x=True, periodic_y=True)
        orc = OF.AdvectionOracle(p, mesh, vm.cv, vm.cw)
        xy = mesh.upts_xy(orc.fr.upts)
        prof = 1 + 0.1 * np.sin(2 * np.pi * xy[:, :, 0])
        f = prof[:, None, :] * np.ones(vm.n)[None, :, None]
        exact = -vm.cv[0][None, :, None] * (0.2 * np.pi * np.cos(2 * np.pi * xy[:, :, 0]))[:, None, :]
        errs.append(np.abs(orc.rhs(f) - exact).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > p - 0.3), (errs, rates)


def _xchannel_rate(vm, left, right, g, p=2):
    """Per-velocity rate of change of the domain total, sum_e sum_p W_p |J| rhs,
    for the uniform state g in a y-periodic channel [0, 1.5] x [0, 1] with the
    given boundaries at x = 0 (outward normal -x) and x = 1.5."""
    mesh = OF.RectMesh(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=False, periodic_y=True)
    orc = OF.AdvectionOracle(p, mesh, vm.cv, vm.cw, walls={"left": left, "right": right})
    nupts = (p + 1) ** 2
    f = np.broadcast_to(g[None, :, None], (nupts, vm.n, 6)).copy()
    r = orc.rhs(f)
    return np.einsum("p,pve->v", orc.fr.W, r) * orc.djac


FAR = dict(N=8, L=5.0)   # c_x = 0.625 lies between 0 and the inlet's U_x = 0.7


@pytest.mark.parametrize("kind", ["inlet", "inlet-normalshock", "vanish", "outlet", "specular"])
def test_farfield_boundaries_balance_inflow_and_outflow(kind):
    """The discrete FR scheme conserves each velocity's total up to the
    boundary fluxes (interior fluxes cancel), and for the exact upwind flux a
    boundary decides only what enters.  So for a uniform state g with an
    outlet at x = 1.5 (extrapolation: g leaves when c_x > 0 and g enters when
    c_x < 0) the domain total of velocity c changes at the rate
    (what enters at x = 0) - c_x g(c) [c_x > 0] times the channel height 1,
    where what enters through x = 0 is, per boundary:
      inlet (U, f0):   c_x f0 for (c - U).(-x) < 0 i.e. c_x > U_x, else c_x g
                       for the velocities with 0 < c_x <= U_x (extrapolated);
      normal shock f0: c_x f0 for every c_x > 0;
      vanish:          nothing;
      outlet:          c_x g for c_x > 0;
      specular:        c_x g(-c_x, c_y, c_z) for c_x > 0 (the mirror image).
    A reversed inequality, the wrong velocity shift or a transposed mirror
    fails this balance."""
    vm = O.VelocityMesh(FAR["N"], FAR["L"], 2, 3)
    cx = vm.cv[0]
    g = OF.maxwellian(vm.cv, 0.9, [0.4, -0.1, 0.05], 1.1)
    f0 = OF.maxwellian(vm.cv, 1.3, [0.7, 0.0, 0.0], 0.8)
    U = np.array([0.7, 0.0, 0.0])
    mirror = OF.maxwellian(vm.cv * np.array([[-1.0], [1.0], [1.0]]), 0.9, [0.4, -0.1, 0.05], 1.1)
    pos = cx > 0
    inflow = {
        "inlet": np.where(cx > U[0], cx * f0, np.where(pos, cx * g, 0.0)),
        "inlet-normalshock": np.where(pos, cx * f0, 0.0),
        "vanish": np.zeros_like(cx),
        "outlet": np.where(pos, cx * g, 0.0),
        "specular": np.where(pos, cx * mirror, 0.0),
    }[kind]
    left = {"inlet": ("inlet", U, f0), "inlet-normalshock": ("inlet-normalshock", f0), "vanish": ("vanish",),
            "outlet": ("outlet",), "specular": ("specular",)}[kind]
    assert np.any((cx > 0) & (cx <= U[0]))
    rate = _xchannel_rate(vm, left, ("outlet",), g)
    want = 1.0 * (inflow - np.where(pos, cx * g, 0.0))
    scale = np.abs(cx * g).max()
    assert np.abs(rate - want).max() < 1e-12 * scale


def test_farfield_free_stream():
    """A uniform state equal to the inflow Maxwellian is steady between an
    inlet and an outlet (and between two normal-shock inlets)."""
    vm = O.VelocityMesh(FAR["N"], FAR["L"], 2, 3)
    U = np.array([0.7, 0.0, 0.0])
    f0 = OF.maxwellian(vm.cv, 1.3, U, 0.8)
    mesh = OF.RectMesh(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=False, periodic_y=True)
    f = np.broadcast_to(f0[None, :, None], (9, vm.n, 6)).copy()
    for walls in ({"left": ("inlet", U, f0), "right": ("outlet",)},
                  {"left": ("inlet-normalshock", f0), "right": ("inlet-normalshock", f0)}):
        r = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw, walls=walls).rhs(f)
        assert np.abs(r).max() < 1e-12 * np.abs(f0).max()


def test_specular_keeps_mirror_symmetric_states():
    """A state even in c_x between specular walls at x = 0 and x = 1.5 has a
    ghost equal to itself: with a uniform state the right-hand side is zero."""
    vm = O.VelocityMesh(FAR["N"], FAR["L"], 2, 3)
    g = OF.maxwellian(vm.cv, 0.9, [0.0, 0.2, 0.0], 1.1)
    mesh = OF.RectMesh(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=False, periodic_y=True)
    orc = OF.AdvectionOracle(2, mesh, vm.cv, vm.cw, walls={"left": ("specular",), "right": ("specular",)})
    f = np.broadcast_to(g[None, :, None], (9, vm.n, 6)).copy()
    assert np.abs(orc.rhs(f)).max() < 1e-12 * np.abs(g).max()


# ---------------------------------------------------------------------------
# triangles
# ---------------------------------------------------------------------------
def _tri_verts(nx, ny, lx, ly):
    """Rectangles split along their rising diagonal, vertices in the
    reference order (-1,-1), (1,-1), (-1,1) of each triangle."""
    xs, ys = np.linspace(0, lx, nx + 1), np.linspace(0, ly, ny + 1)
    out = []
    for iy in range(ny):
        for ix in range(nx):
            a, b = (xs[ix], ys[iy]), (xs[ix + 1], ys[iy])
            c, d = (xs[ix + 1], ys[iy + 1]), (xs[ix], ys[iy + 1])
            out += [[a, b, d], [c, d, b]]
    return np.array(out)


def _ybound(ly):
    return lambda m: "bottom" if abs(m[1]) < 1e-9 else ("top" if abs(m[1] - ly) < 1e-9 else "side")


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tri_interpolation_derivative_and_weights(p):
    t = OF.TriFR(p)
    rng = np.random.default_rng(p)
    c = {e: rng.normal() for e in t.exps}
    poly = lambda x, y: sum(v * x ** i * y ** j for (i, j), v in c.items())
    dpx = lambda x, y: sum(v * i * x ** max(i - 1, 0) * y ** j for (i, j), v in c.items())
    dpy = lambda x, y: sum(v * j * x ** i * y ** max(j - 1, 0) for (i, j), v in c.items())
    u = poly(t.upts[:, 0], t.upts[:, 1])
    assert np.allclose(t.basis_at(t.fpts) @ u, poly(t.fpts[:, 0], t.fpts[:, 1]), atol=1e-12)
    g = t.grad_at(t.upts)
    assert np.allclose(g[0] @ u, dpx(t.upts[:, 0], t.upts[:, 1]), atol=1e-10)
    assert np.allclose(g[1] @ u, dpy(t.upts[:, 0], t.upts[:, 1]), atol=1e-10)
    # the integrals of the nodal basis integrate P_p exactly (area 2)
    assert abs(t.W.sum() - 2.0) < 1e-13
    assert abs(t.W @ t.upts[:, 0] + 2.0 / 3.0) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tri_lift_discrete_divergence(p):
    """Gauss-Green in discrete form: for u in P_p, the FR divergence of the
    transformed flux (a, b) u, corrected with the face fluxes of u itself, is
    the exact divergence; and the lift of a face flux integrates to its face
    integral: sum_u W_u M3[u, (f, k)] = w_k."""
    t = OF.TriFR(p)
    M3 = t.lift()
    n = p + 1
    for f in range(3):
        for k in range(n):
            assert abs(t.W @ M3[:, f * n + k] - t.w[k]) < 1e-12
    rng = np.random.default_rng(10 + p)
    c = {e: rng.normal() for e in t.exps}
    poly = lambda x, y: sum(v * x ** i * y ** j for (i, j), v in c.items())
    u = poly(t.upts[:, 0], t.upts[:, 1])
    a, b = 0.7, -1.3
    D = t.grad_at(t.upts)
    uf = t.basis_at(t.fpts) @ u
    div = a * D[0] @ u + b * D[1] @ u
    fn = (t.fnorm[:, 0] * a + t.fnorm[:, 1] * b) * uf
    # with the common flux equal to the interior one the correction vanishes
    assert np.allclose(div + M3 @ (fn - fn), div)
    # total: sum W div = face integral of n.(a, b) u (face measure ds)
    tot = t.W @ div
    zf, wf = np.polynomial.legendre.leggauss(n + 2)
    face = 0.0
    for proj, nrm in t.faces:
        xs, ys = proj(zf)
        face += np.sum(wf * (nrm[0] * a + nrm[1] * b) * poly(xs, ys))
    assert abs(tot - face) < 1e-11 * max(1.0, abs(face))


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tri_product_operators_match_oracle(p):
    from dgfs_hip.shapes import get_shape
    s = get_shape("tri", p)
    t = OF.TriFR(p)
    assert np.allclose(s.upts, t.upts, atol=2e-15)
    assert np.allclose(s.fpts, t.fpts, atol=1e-15)
    assert np.allclose(s.m0, t.basis_at(t.fpts), atol=1e-12)
    D = t.grad_at(t.upts)
    assert np.allclose(s.m1, np.concatenate([D[0], D[1]], axis=1), atol=1e-10)
    assert np.allclose(s.m3, t.lift(), atol=1e-11)


def test_tri_free_stream_and_conservation():
    vm = _vm()
    verts = _tri_verts(3, 2, 1.5, 1.0)
    orc = OF.TriMeshAdvectionOracle(2, verts, vm.cv, vm.cw, period=1.5,
                                    walls={"bottom": ("specular",), "top": ("specular",)},
                                    boundary_name=_ybound(1.0))
    g = OF.maxwellian(vm.cv, 1.0, [0.3, 0.0, 0.1], 0.9)     # even in c_y
    f = np.broadcast_to(g[None, :, None], (6, vm.n, len(verts))).copy()
    assert np.abs(orc.rhs(f)).max() < 1e-12 * g.max()
    rng = np.random.default_rng(3)
    f = f * (1 + 0.1 * rng.random(f.shape))
    r = orc.rhs(f)
    # specular walls reflect every particle: each mirror pair of velocities
    # keeps its combined total
    tot = np.einsum("p,pve,e->v", orc.tri.W, r, orc.djac)
    mirror = OF.AdvectionOracle._farfield_ghost(orc, ("specular",), np.array([0.0, 1.0]), np.arange(vm.n)[None].astype(float))[0].astype(int)
    assert np.abs(tot + tot[mirror]).max() < 1e-11 * np.abs(r).max()


@pytest.mark.parametrize("p", [1, 2])
def test_tri_advection_converges(p):
    """f = 1 + 0.1 sin(2 pi x), uniform in y and even in c_y between specular
    walls: the rhs is -c_x 0.2 pi cos(2 pi x), with error falling at order p."""
    vm = O.VelocityMesh(4, 3.0, 2, 3)
    errs = []
    for nx in (4, 8, 16):
        verts = _tri_verts(nx, 2, 1.0, 0.5)
        orc = OF.TriMeshAdvectionOracle(p, verts, vm.cv, vm.cw, period=1.0,
                                        walls={"bottom": ("specular",), "top": ("specular",)},
                                        boundary_name=_ybound(0.5))
        lin = OF.TriMeshAdvectionOracle._lin(orc.tri.upts)
        xy = np.einsum("ua,ead->ued", lin, verts)
        x = xy[:, :, 0]
        f = (1 + 0.1 * np.sin(2 * np.pi * x))[:, None, :] * np.ones(vm.n)[None, :, None]
        exact = -vm.cv[0][None, :, None] * (0.2 * np.pi * np.cos(2 * np.pi * x))[:, None, :]
        errs.append(float(np.abs(orc.rhs(f) - exact).max()))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > p - 0.3), (errs, rates)


def test_opmat_expression_parser():
    """'*' binds tighter than '+'/'-', unary minus and parentheses work, and
    anything but M<n>, + - * ( ) is refused (the syntax of shapes.py:83-91)."""
    from dgfs_hip.shapes import eval_opmat
    rng = np.random.default_rng(0)
    mats = {f"M{i}": rng.normal(size=(3, 3)) for i in range(4)}
    get = mats.__getitem__
    M0, M1, M2, M3 = (mats[f"M{i}"] for i in range(4))
    assert np.allclose(eval_opmat("M1 - M3*M2", get), M1 - M3 @ M2)
    assert np.allclose(eval_opmat("(M1 - M3*M2)*M0", get), (M1 - M3 @ M2) @ M0)
    assert np.allclose(eval_opmat("-M0 + M1*M2*M3", get), -M0 + M1 @ M2 @ M3)
    assert np.allclose(eval_opmat("M0 - M1 - M2", get), M0 - M1 - M2)
    for bad in ("M1 ** M2", "M1 + __import__", "(M1 - M2", "M1 M2", "M1 +", "M1/M2"):
        with pytest.raises((ValueError, KeyError)):
            eval_opmat(bad, get)
