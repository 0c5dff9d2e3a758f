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
