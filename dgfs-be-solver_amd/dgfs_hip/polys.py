"""Orthonormal polynomial bases on the reference line, quadrilateral and
triangle (mirrors frfs/polys.py:12-208: Jacobi recurrences, tensor Legendre
on quads, Dubiner/Koornwinder on triangles; nodal bases by solving with the
Vandermonde matrix)."""
import math

import numpy as np


def jacobi(n, a, b, z):
    """[P_0 .. P_n] of the Jacobi polynomials P^(a,b) at z (polys.py:12-31)."""
    z = np.asarray(z, dtype=float)
    j = [np.ones_like(z)]
    if n >= 1:
        j.append(((a + b + 2) * z + a - b) / 2)
    if n >= 2:
        apb, bbmaa = a + b, b * b - a * a
        for q in range(2, n + 1):
            qapbpq, apbp2q = q * (apb + q), apb + 2 * q
            apbp2qm1, apbp2qm2 = apbp2q - 1, apbp2q - 2
            aq = apbp2q * apbp2qm1 / (2 * qapbpq)
            bq = apbp2qm1 * bbmaa / (2 * qapbpq * apbp2qm2)
            cq = apbp2q * (a + q - 1) * (b + q - 1) / (qapbpq * apbp2qm2)
            j.append((aq * z - bq) * j[-1] - cq * j[-2])
    return j


def jacobi_diff(n, a, b, z):
    z = np.asarray(z, dtype=float)
    dj = [np.zeros_like(z)]
    if n >= 1:
        dj.extend(jp * (i + a + b + 2) / 2 for i, jp in enumerate(jacobi(n - 1, a + 1, b + 1, z)))
    return dj


class BasePolyBasis:
    name = None

    def __init__(self, order, pts):
        self.order = order
        self.pts = np.asarray(pts, dtype=float)

    @property
    def vdm(self):
        return self.ortho_basis_at(self.pts)

    def nodal_basis_at(self, epts):
        """L[p][j]: the nodal basis of self.pts evaluated at epts."""
        return np.linalg.solve(self.vdm, self.ortho_basis_at(epts)).T

    def jac_nodal_basis_at(self, epts):
        """dL[d][j][p]: derivative along reference axis d of nodal basis j at p."""
        J = self.jac_ortho_basis_at(epts)  # (ndims, nbasis, npts)
        return np.stack([np.linalg.solve(self.vdm, J[d]) for d in range(J.shape[0])])


class LinePolyBasis(BasePolyBasis):
    name = "line"

    def ortho_basis_at(self, pts):
        p = np.asarray(pts, dtype=float).reshape(-1)
        jp = jacobi(self.order - 1, 0, 0, p)
        return np.array([math.sqrt(i + 0.5) * v for i, v in enumerate(jp)])

    def jac_ortho_basis_at(self, pts):
        p = np.asarray(pts, dtype=float).reshape(-1)
        djp = jacobi_diff(self.order - 1, 0, 0, p)
        return np.array([[math.sqrt(i + 0.5) * v for i, v in enumerate(djp)]])

    @property
    def degrees(self):
        return [(i,) for i in range(self.order)]


class QuadPolyBasis(BasePolyBasis):
    name = "quad"

    def _pab(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=float))
        p, q = pts[:, 0], pts[:, 1]
        sk = [math.sqrt(k + 0.5) for k in range(self.order)]
        pa = [c * jp for c, jp in zip(sk, jacobi(self.order - 1, 0, 0, p))]
        pb = [c * jp for c, jp in zip(sk, jacobi(self.order - 1, 0, 0, q))]
        dpa = [c * jp for c, jp in zip(sk, jacobi_diff(self.order - 1, 0, 0, p))]
        dpb = [c * jp for c, jp in zip(sk, jacobi_diff(self.order - 1, 0, 0, q))]
        return pa, pb, dpa, dpb

    def ortho_basis_at(self, pts):
        pa, pb, _, _ = self._pab(pts)
        return np.array([pi * pj for pi in pa for pj in pb])

    def jac_ortho_basis_at(self, pts):
        pa, pb, dpa, dpb = self._pab(pts)
        d0 = [dpi * pj for pi, dpi in zip(pa, dpa) for pj, dpj in zip(pb, dpb)]
        d1 = [pi * dpj for pi, dpi in zip(pa, dpa) for pj, dpj in zip(pb, dpb)]
        return np.array([d0, d1])

    @property
    def degrees(self):
        return [(i, j) for i in range(self.order) for j in range(self.order)]


class TriPolyBasis(BasePolyBasis):
    name = "tri"

    def _ab(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=float))
        p, q = pts[:, 0], pts[:, 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            a = np.where(q != 1, 2 * (1 + p) / (1 - q) - 1, -1.0)
        return a, q

    def ortho_basis_at(self, pts):
        a, b = self._ab(pts)
        ob = []
        for i, pi in enumerate(jacobi(self.order - 1, 0, 0, a)):
            pa = pi * (1 - b) ** i
            for j, pj in enumerate(jacobi(self.order - i - 1, 2 * i + 1, 0, b)):
                cij = math.sqrt((2 * i + 1) * (2 * i + 2 * j + 2)) / 2 ** (i + 1)
                ob.append(cij * pa * pj)
        return np.array(ob)

    def jac_ortho_basis_at(self, pts):
        a, b = self._ab(pts)
        f = jacobi(self.order - 1, 0, 0, a)
        df = jacobi_diff(self.order - 1, 0, 0, a)
        d0, d1 = [], []
        for i, (fi, dfi) in enumerate(zip(f, df)):
            g = jacobi(self.order - i - 1, 2 * i + 1, 0, b)
            dg = jacobi_diff(self.order - i - 1, 2 * i + 1, 0, b)
            for j, (gj, dgj) in enumerate(zip(g, dg)):
                cij = math.sqrt((2 * i + 1) * (2 * i + 2 * j + 2)) / 2 ** (i + 1)
                tmp = (1 - b) ** (i - 1) if i > 0 else np.ones_like(b)
                d0.append(cij * 2 * tmp * dfi * gj)
                d1.append(cij * (tmp * (-i * fi + (1 + a) * dfi) * gj + (1 - b) ** i * fi * dgj))
        return np.array([d0, d1])

    @property
    def degrees(self):
        return [(i, j) for i in range(self.order) for j in range(self.order - i)]


def get_polybasis(name, order, pts=()):
    return {"line": LinePolyBasis, "quad": QuadPolyBasis, "tri": TriPolyBasis}[name](order, pts)
