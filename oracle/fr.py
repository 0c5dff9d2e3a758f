"""FP64 CPU oracle for the flux-reconstruction advection of the DGFS solver.

TEST INFRASTRUCTURE ONLY (used by tests/ as the checker; never imported by
the product).  It restates the advection right-hand side of
frfs/solvers/dgfs/system.py:140-177 for structured rectangles of
quadrilaterals, building every operator from the mathematics in a way that
differs from the product's code path:

* nodal basis: tensor products of 1D Lagrange polynomials on the Gauss
  points (the product inverts a Legendre Vandermonde matrix, polys.py);
* correction lift M3 (shapes.py:107-118, 243-289): the DG lift with the
  (p+1)-point Gauss-Lobatto face rule, obtained from the diagonal Gauss mass
  matrix, M3[p][k] = (1/W_p) sum_i w_i l_k(s_i) L_p(face(s_i));
* geometry: closed-form metric terms of an axis-aligned rectangle
  (smats = |J| J^-1 = diag(hy/2, hx/2), |J| = hx hy / 4);
* fluxes: the exact upwind flux of rusanov.mako:47-49, the transformed flux
  of tflux.mako/flux.mako, the diffuse wall of bcs/dgfs-wall-diffuse*.mako
  and negdivconf.mako, assembled element by element.

How it is pinned (tests/test_oracle_fr.py): exact interpolation and
differentiation of polynomials of degree <= p; the lift identity
sum_p W_p M3[p][k] q(x_p) = face integral of l_k q for q of degree <= p;
free-stream preservation on periodic meshes; discrete conservation
(sum of W |J| rhs equals the boundary flux); convergence of smooth periodic
advection at order p+1.
"""
from __future__ import annotations

import numpy as np


def gauss_legendre(n):
    return np.polynomial.legendre.leggauss(n)


def gauss_lobatto(n):
    # interior nodes are the roots of P'_{n-1}; weights 2/(n(n-1) P_{n-1}^2)
    if n == 2:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    c = np.zeros(n)
    c[-1] = 1.0
    inner = np.polynomial.legendre.Legendre(c).deriv().roots()
    z = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    Pn = np.polynomial.legendre.legval(z, c)
    w = 2.0 / (n * (n - 1) * Pn ** 2)
    return z, w


def lagrange(nodes, x):
    """L[j](x) for the 1D Lagrange basis on `nodes`; x array -> (len(x), n)."""
    x = np.atleast_1d(np.asarray(x, dtype=float))
    n = len(nodes)
    out = np.ones((len(x), n))
    for j in range(n):
        for m in range(n):
            if m != j:
                out[:, j] *= (x - nodes[m]) / (nodes[j] - nodes[m])
    return out


def lagrange_deriv(nodes, x):
    x = np.atleast_1d(np.asarray(x, dtype=float))
    n = len(nodes)
    out = np.zeros((len(x), n))
    for j in range(n):
        for k in range(n):
            if k == j:
                continue
            term = np.ones(len(x)) / (nodes[j] - nodes[k])
            for m in range(n):
                if m != j and m != k:
                    term *= (x - nodes[m]) / (nodes[j] - nodes[m])
            out[:, j] += term
    return out


class QuadFR:
    """Operators of one quadrilateral of order p with Gauss solution and flux
    points.  Solution points are x-fastest tensor points; flux points run face
    by face (bottom, right, top, left), ascending in the face coordinate."""

    faces = [(lambda s: (s, -1.0 + 0 * s), (0.0, -1.0)), (lambda s: (1.0 + 0 * s, s), (1.0, 0.0)),
             (lambda s: (s, 1.0 + 0 * s), (0.0, 1.0)), (lambda s: (-1.0 + 0 * s, s), (-1.0, 0.0))]

    def __init__(self, p):
        self.p = p
        n = p + 1
        self.z, self.w = gauss_legendre(n)
        self.nupts = n * n
        self.upts = np.array([(self.z[i], self.z[j]) for j in range(n) for i in range(n)])
        self.W = np.array([self.w[i] * self.w[j] for j in range(n) for i in range(n)])
        fp, fn = [], []
        for proj, nrm in self.faces:
            xs, ys = proj(self.z)
            fp.append(np.stack([xs, ys], axis=1))
            fn.append(np.tile(nrm, (n, 1)))
        self.fpts = np.vstack(fp)
        self.fnorm = np.vstack(fn)
        self.nfpts = len(self.fpts)

    def basis_at(self, pts):
        """(npts, nupts) nodal basis values."""
        Lx = lagrange(self.z, pts[:, 0])
        Ly = lagrange(self.z, pts[:, 1])
        n = self.p + 1
        return np.stack([Lx[:, i] * Ly[:, j] for j in range(n) for i in range(n)], axis=1)

    def grad_at(self, pts):
        """(2, npts, nupts) derivatives along xi, eta."""
        Lx, Ly = lagrange(self.z, pts[:, 0]), lagrange(self.z, pts[:, 1])
        dLx, dLy = lagrange_deriv(self.z, pts[:, 0]), lagrange_deriv(self.z, pts[:, 1])
        n = self.p + 1
        gx = np.stack([dLx[:, i] * Ly[:, j] for j in range(n) for i in range(n)], axis=1)
        gy = np.stack([Lx[:, i] * dLy[:, j] for j in range(n) for i in range(n)], axis=1)
        return np.stack([gx, gy])

    def lift(self):
        """M3 (nupts, nfpts): DG lift with the (p+1)-point GLL face rule."""
        zq, wq = gauss_lobatto(self.p + 1)
        n = self.p + 1
        M3 = np.zeros((self.nupts, self.nfpts))
        for f, (proj, _) in enumerate(self.faces):
            lk = lagrange(self.z, zq)                       # face nodal basis at GLL points
            xs, ys = proj(zq)
            Lp = self.basis_at(np.stack([xs, ys], axis=1))  # (nq, nupts)
            for k in range(n):
                M3[:, f * n + k] = (wq * lk[:, k]) @ Lp / self.W
        return M3


class RectMesh:
    """nx x ny rectangles of [x0,x1]x[y0,y1]; element e = ix + nx*iy."""

    def __init__(self, nx, ny, x0, x1, y0, y1, periodic_x, periodic_y=False):
        self.nx, self.ny = nx, ny
        self.hx, self.hy = (x1 - x0) / nx, (y1 - y0) / ny
        self.x0, self.y0 = x0, y0
        self.periodic_x = periodic_x
        self.periodic_y = periodic_y

    def neighbour(self, e, f):
        """(e', f') across face f of element e, or None on a boundary."""
        ix, iy = e % self.nx, e // self.nx
        if f == 1:
            if ix + 1 < self.nx:
                return e + 1, 3
            return (e - ix, 3) if self.periodic_x else None
        if f == 3:
            if ix > 0:
                return e - 1, 1
            return (e + self.nx - 1, 1) if self.periodic_x else None
        if f == 2:
            if iy + 1 < self.ny:
                return e + self.nx, 0
            return (ix, 0) if self.periodic_y else None
        if iy > 0:
            return e - self.nx, 2
        return (ix + self.nx * (self.ny - 1), 2) if self.periodic_y else None

    def upts_xy(self, upts):
        """Physical coordinates (nupts, neles, 2) of reference points."""
        out = np.empty((len(upts), self.nx * self.ny, 2))
        for e in range(self.nx * self.ny):
            ix, iy = e % self.nx, e // self.nx
            out[:, e, 0] = self.x0 + (ix + (upts[:, 0] + 1) / 2) * self.hx
            out[:, e, 1] = self.y0 + (iy + (upts[:, 1] + 1) / 2) * self.hy
        return out

    def boundary_name(self, e, f):
        return {0: "bottom", 1: "right", 2: "top", 3: "left"}[f]


def maxwellian(cv, rho, u, T):
    u = np.asarray(u, dtype=float).reshape(3, 1)
    return rho / (np.pi * T) ** 1.5 * np.exp(-np.sum((cv - u) ** 2, axis=0) / T)


class AdvectionOracle:
    """FP64 advection rhs of a DGFS state f[upt][v][e] on a RectMesh.

    walls: {name: (U (3,), f_w (vsize,))} diffuse walls, or {name: fn} for
    a wall whose velocity and temperature vary along it
    (bcs/dgfs-wall-diffuse-cyl.mako): fn(xy (n, 2)) -> (U (n, 3), T (n,)),
    emitting exp(-|c - U|^2 / T) with the same flux balance, or
    {name: (kind, *args)} for the far-field and specular boundaries:
      ("inlet", U, f0)          bcs/dgfs-inlet.mako: f0 where (c - U).n < 0
      ("inlet-normalshock", f0) bcs/dgfs-inlet-normalshock.mako: f0 where c.n < 0
      ("outlet",)               bcs/dgfs-outlet.mako: the interior state
      ("vanish",)               bcs/dgfs-vanish.mako: zero where c.n < 0
      ("specular",)             bcs/dgfs-wall-specular*.mako: the interior
                                state at the mirrored velocity c - 2 (c.n) n
    (n the outward unit normal; the dot products run over the mesh's two
    dimensions).  Faces without a neighbour must be listed."""

    def __init__(self, p, mesh: RectMesh, cv, cw, walls=None):
        self.fr = fr = QuadFR(p)
        self.mesh = mesh
        self.cv = cv
        self.cw = cw
        self.walls = walls or {}
        self.B = fr.basis_at(fr.fpts)            # M0
        self.D = fr.grad_at(fr.upts)             # (2, nupts, nupts)
        self.M3 = fr.lift()
        self.sm = np.array([mesh.hy / 2, mesh.hx / 2])   # diagonal smats
        self.djac = mesh.hx * mesh.hy / 4

    def rhs(self, f):
        fr, mesh, cv = self.fr, self.mesh, self.cv
        nupts, vsize, neles = f.shape
        n = fr.p + 1
        uf = np.einsum("fu,uve->fve", self.B, f)                  # flux-point states
        # transformed flux divergence at solution points
        div = np.zeros_like(f)
        for d in range(2):
            Fd = self.sm[d] * cv[d][None, :, None] * f
            div += np.einsum("pu,uve->pve", self.D[d], Fd)
        # interior normal transformed flux at flux points (reference normals)
        Fn_int = np.zeros_like(uf)
        for d in range(2):
            Fn_int += fr.fnorm[:, d][:, None, None] * self.sm[d] * cv[d][None, :, None] * uf
        # common flux
        Fn_com = np.zeros_like(uf)
        for e in range(neles):
            for fc in range(4):
                rows = np.arange(fc * n, (fc + 1) * n)
                nref = np.array(fr.faces[fc][1])
                pn = nref * self.sm                                   # physical normal * mag
                mag = np.linalg.norm(pn)
                nrm = pn / mag
                nc = nrm[0] * cv[0] + nrm[1] * cv[1]
                nb = mesh.neighbour(e, fc)
                ul = uf[rows, :, e]
                if nb is not None:
                    e2, f2 = nb
                    # faces are parametrised bottom-to-top / left-to-right on
                    # both sides, so neighbouring flux points share the order
                    rows2 = np.arange(f2 * n, (f2 + 1) * n)
                    ur = uf[rows2, :, e2]
                else:
                    wall = self.walls[mesh.boundary_name(e, fc)]
                    if isinstance(wall, tuple) and isinstance(wall[0], str):
                        ur = self._farfield_ghost(wall, nrm, ul)
                    elif callable(wall):
                        xy = mesh.upts_xy(fr.fpts[rows])[:, e]
                        Up, Tp = wall(xy)
                        ur = np.empty_like(ul)
                        for r in range(n):
                            fw = np.exp(-np.sum((cv - Up[r][:, None]) ** 2, axis=0) / Tp[r])
                            ur[r] = self._diffuse_ghost(nrm, Up[r], fw, ul[r:r + 1])[0]
                    else:
                        U, fw = wall
                        ur = self._diffuse_ghost(nrm, U, fw, ul)
                fn = 0.5 * nc * (ul + ur) + 0.5 * np.abs(nc) * (ul - ur)
                Fn_com[rows, :, e] = mag * fn
        corr = np.einsum("pf,fve->pve", self.M3, Fn_com - Fn_int)
        return -(div + corr) / self.djac

    def _diffuse_ghost(self, nrm, U, fw, ul):
        """Ghost state of a diffuse wall for wall-side states ul (rows, vsize):
        the wall Maxwellian on incoming velocities, scaled so that the
        normal mass flux through every point vanishes."""
        cv = self.cv
        fac = nrm[0] * (cv[0] - U[0]) + nrm[1] * (cv[1] - U[1])
        num = np.sum(np.where(fac >= 0, fac * ul * self.cw, 0.0), axis=1)
        den = np.sum(np.where(fac < 0, fac * fw * self.cw, 0.0))
        return np.where(fac[None, :] < 0, fw[None, :] * (-num[:, None] / den), ul)

    def _farfield_ghost(self, wall, nrm, ul):
        """Ghost states of the inflow/outflow and specular boundaries for the
        wall-side states ul (rows, vsize); nrm is the outward unit normal."""
        cv = self.cv
        kind = wall[0]
        cn = nrm[0] * cv[0] + nrm[1] * cv[1]
        if kind == "outlet":
            return ul.copy()
        if kind == "vanish":
            return np.where(cn[None, :] < 0, 0.0, ul)
        if kind == "inlet-normalshock":
            f0 = wall[1]
            return np.where(cn[None, :] < 0, f0[None, :], ul)
        if kind == "inlet":
            U, f0 = wall[1], wall[2]
            fac = nrm[0] * (cv[0] - U[0]) + nrm[1] * (cv[1] - U[1])
            return np.where(fac[None, :] < 0, f0[None, :], ul)
        if kind == "specular":
            # index of the mirror image of every velocity (the grid is
            # symmetric about the origin, so mirrors of grid points along an
            # axis-aligned normal are grid points)
            n3 = np.array([nrm[0], nrm[1], 0.0])
            cr = cv - 2.0 * (n3 @ cv)[None, :] * n3[:, None]
            key = {tuple(np.round(cv[:, j], 9)): j for j in range(cv.shape[1])}
            mirror = np.array([key[tuple(np.round(cr[:, j], 9))] for j in range(cv.shape[1])])
            return ul[:, mirror]
        raise ValueError(f"unknown boundary kind {kind!r}")


# ---------------------------------------------------------------------------
# Triangles (frfs/shapes.py:404-425 TriShape; frfs/polys.py:135-181; the lift
# of shapes.py:243-289).  Built here from monomials and the definition of the
# DG lift, independently of the product's orthonormal-basis construction.
# ---------------------------------------------------------------------------
# Williams-Shunn-Jameson solution points on the reference triangle
# (-1,-1), (1,-1), (-1,1): the data of frfs/quadrules/tri/williams-shunn-n{3,6,10}-*.txt
WS_TRI = {
    1: [(-0.666666666666666, 0.333333333333332), (0.333333333333332, -0.666666666666666),
        (-0.666666666666666, -0.666666666666666)],
    2: [(-0.81684757298044, 0.63369514596088), (0.63369514596088, -0.81684757298044),
        (-0.81684757298044, -0.81684757298044), (-0.108103018168072, -0.783793963663856),
        (-0.783793963663856, -0.108103018168072), (-0.108103018168072, -0.108103018168072)],
    3: [(-0.333333333333333, -0.333333333333333), (-0.888871894660414, 0.777743789320828),
        (0.777743789320828, -0.888871894660414), (-0.888871894660414, -0.888871894660414),
        (0.268421495491446, -0.408932576528214), (-0.859488918963232, -0.408932576528214),
        (-0.408932576528214, -0.859488918963232), (0.268421495491446, -0.859488918963232),
        (-0.859488918963232, 0.268421495491446), (-0.408932576528214, 0.268421495491446)],
}


def triangle_rule(n):
    """Collapsed-coordinate Gauss rule with n x n points on the reference
    triangle (exact for degree 2n - 2)."""
    a, wa = np.polynomial.legendre.leggauss(n)
    A, B = np.meshgrid(a, a, indexing="ij")
    W = np.outer(wa, wa) * (1.0 - B) / 2.0
    xi = (1.0 + A) * (1.0 - B) / 2.0 - 1.0
    return np.stack([xi.ravel(), B.ravel()], axis=1), W.ravel()


class TriFR:
    """Operators of one triangle of order p.  Faces (type line):
    0: (s, -1), normal (0, -1); 1: (-s, s), normal (1, 1); 2: (-1, s),
    normal (-1, 0); flux points are the p+1 Gauss points in s, ascending."""

    faces = [(lambda s: (s, -1.0 + 0 * s), (0.0, -1.0)), (lambda s: (-s, s), (1.0, 1.0)),
             (lambda s: (-1.0 + 0 * s, s), (-1.0, 0.0))]

    def __init__(self, p, upts=None):
        self.p = p
        self.exps = [(i, j) for i in range(p + 1) for j in range(p + 1 - i)]
        self.upts = np.array(upts if upts is not None else WS_TRI[p], dtype=float)
        self.nupts = len(self.upts)
        assert self.nupts == len(self.exps)
        self.Vinv = np.linalg.inv(self._mono(self.upts))
        self.z, self.w = gauss_legendre(p + 1)
        fp, fn = [], []
        for proj, nrm in self.faces:
            xs, ys = proj(self.z)
            fp.append(np.stack([xs, ys], axis=1))
            fn.append(np.tile(nrm, (p + 1, 1)))
        self.fpts = np.vstack(fp)
        self.fnorm = np.vstack(fn)
        self.nfpts = len(self.fpts)
        qp, qw = triangle_rule(p + 3)
        B = self.basis_at(qp)
        self.W = qw @ B                          # integrals of the nodal basis

    def _mono(self, pts):
        x, y = pts[:, 0], pts[:, 1]
        return np.stack([x ** i * y ** j for i, j in self.exps], axis=1)

    def _dmono(self, pts):
        x, y = pts[:, 0], pts[:, 1]
        dx = np.stack([i * x ** max(i - 1, 0) * y ** j for i, j in self.exps], axis=1)
        dy = np.stack([j * x ** i * y ** max(j - 1, 0) for i, j in self.exps], axis=1)
        return dx, dy

    def basis_at(self, pts):
        return self._mono(np.atleast_2d(pts)) @ self.Vinv

    def grad_at(self, pts):
        dx, dy = self._dmono(np.atleast_2d(pts))
        return np.stack([dx @ self.Vinv, dy @ self.Vinv])

    def lift(self):
        """M3 (nupts, nfpts): the DG lift g_k in P_p with
        int_T g_k q = sum_i wq_i l_k(s_i) q(face(s_i)) for every q in P_p,
        face integrals by the (p+1)-point Gauss-Lobatto rule in s."""
        qp, qw = triangle_rule(self.p + 3)
        Mq = self._mono(qp)
        mass = (Mq * qw[:, None]).T @ Mq
        zq, wq = gauss_lobatto(self.p + 1)
        lk = lagrange(self.z, zq)
        n = self.p + 1
        M3 = np.zeros((self.nupts, self.nfpts))
        Mu = self._mono(self.upts)
        for f, (proj, _) in enumerate(self.faces):
            xs, ys = proj(zq)
            Mf = self._mono(np.stack([xs, ys], axis=1))
            for k in range(n):
                rhs = (wq * lk[:, k]) @ Mf
                M3[:, f * n + k] = Mu @ np.linalg.solve(mass, rhs)
        return M3


class TriMeshAdvectionOracle:
    """FP64 advection rhs on straight-sided triangles.

    verts: (neles, 3, 2) vertex coordinates in the reference order
    (-1,-1), (1,-1), (-1,1); period: x-period of the mesh (None: not
    periodic); walls: {name: (U, f_w) | (kind, *args)} as AdvectionOracle,
    keyed by boundary_name(midpoint) for faces without a neighbour."""

    def __init__(self, p, verts, cv, cw, walls=None, period=None, boundary_name=None):
        self.tri = tri = TriFR(p)
        self.verts = np.asarray(verts, dtype=float)
        self.cv, self.cw = cv, cw
        self.walls = walls or {}
        self.period = period
        self.boundary_name = boundary_name
        self.B = tri.basis_at(tri.fpts)
        self.D = tri.grad_at(tri.upts)
        self.M3 = tri.lift()
        ne = len(self.verts)
        v0, v1, v2 = self.verts[:, 0], self.verts[:, 1], self.verts[:, 2]
        # x(xi, eta) = v0 + (v1 - v0)(xi + 1)/2 + (v2 - v0)(eta + 1)/2
        self.J = np.stack([(v1 - v0) / 2.0, (v2 - v0) / 2.0], axis=2)      # (ne, phys, ref)
        self.djac = self.J[:, 0, 0] * self.J[:, 1, 1] - self.J[:, 0, 1] * self.J[:, 1, 0]
        # smats[e, i(ref), k(phys)] = djac * (J^-1)[i, k]
        self.smats = np.empty((ne, 2, 2))
        self.smats[:, 0, 0], self.smats[:, 0, 1] = self.J[:, 1, 1], -self.J[:, 0, 1]
        self.smats[:, 1, 0], self.smats[:, 1, 1] = -self.J[:, 1, 0], self.J[:, 0, 0]
        n = p + 1
        self.xf = np.einsum("fa,eda->efd", self._lin(tri.fpts), np.transpose(self.verts, (0, 2, 1)))
        self.match = {}
        mids = {}
        for e in range(ne):
            for f in range(3):
                m = self.xf[e, f * n:(f + 1) * n].mean(axis=0)
                mids[(e, f)] = m
        keys = list(mids)
        for a in keys:
            for b in keys:
                if a == b or a[0] == b[0]:
                    continue
                d = mids[a] - mids[b]
                if self.period is not None:
                    d[0] -= self.period * np.round(d[0] / self.period)
                if np.hypot(*d) < 1e-9:
                    self.match[a] = b

    @staticmethod
    def _lin(pts):
        xi, eta = pts[:, 0], pts[:, 1]
        return np.stack([-(xi + eta) / 2.0, (1.0 + xi) / 2.0, (1.0 + eta) / 2.0], axis=1)

    def rhs(self, f):
        tri, cv = self.tri, self.cv
        nupts, vsize, ne = f.shape
        n = tri.p + 1
        uf = np.einsum("fu,uve->fve", self.B, f)
        out = np.empty_like(f)
        for e in range(ne):
            sm = self.smats[e]
            # transformed flux F_i = sum_k smats[i][k] c_k f
            div = np.zeros((nupts, vsize))
            Fint = np.zeros((tri.nfpts, vsize))
            for i in range(2):
                ci = sm[i, 0] * cv[0] + sm[i, 1] * cv[1]
                div += self.D[i] @ (ci[None, :] * f[:, :, e])
                Fint += tri.fnorm[:, i][:, None] * ci[None, :] * uf[:, :, e]
            Fcom = np.empty_like(Fint)
            for fc in range(3):
                rows = np.arange(fc * n, (fc + 1) * n)
                pn = tri.faces[fc][1] @ sm                  # sum_i nref_i smats[i][k]
                mag = np.linalg.norm(pn)
                nrm = pn / mag
                nc = nrm[0] * cv[0] + nrm[1] * cv[1]
                ul = uf[rows, :, e]
                nb = self.match.get((e, fc))
                if nb is not None:
                    e2, f2 = nb
                    rows2 = np.arange(f2 * n, (f2 + 1) * n)
                    ur = np.empty_like(ul)
                    for r in range(n):
                        d = self.xf[e2, rows2] - self.xf[e, rows[r]]
                        if self.period is not None:
                            d[:, 0] -= self.period * np.round(d[:, 0] / self.period)
                        j = int(np.argmin(np.hypot(d[:, 0], d[:, 1])))
                        ur[r] = uf[rows2[j], :, e2]
                else:
                    name = self.boundary_name(self.xf[e, rows].mean(axis=0))
                    wall = self.walls[name]
                    if isinstance(wall[0], str):
                        ur = AdvectionOracle._farfield_ghost(self, wall, nrm, ul)
                    else:
                        ur = AdvectionOracle._diffuse_ghost(self, nrm, wall[0], wall[1], ul)
                Fcom[rows] = mag * (0.5 * nc * (ul + ur) + 0.5 * np.abs(nc) * (ul - ur))
            out[:, :, e] = -(div + self.M3 @ (Fcom - Fint)) / self.djac[e]
        return out
