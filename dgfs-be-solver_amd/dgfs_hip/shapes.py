"""Reference elements and FR operator matrices (mirrors frfs/shapes.py:25-425).

Operators, as opmat() returns them (shapes.py:83-118):
  M0 = nodal basis of the solution points evaluated at the flux points
  M1 = derivatives of that basis at the solution points, [upt][d*nupts + j]
  M2 = reference face normal times M0, [fpt][d*nupts + j]
  M3 = correction lift: sum_j phi_j(upt) * S[fpt][j], S the face integral of
       face-nodal basis times orthonormal basis on a (p+1)-point
       Gauss-Lobatto-Legendre face rule (gbasis_coeffs, shapes.py:243-285)
  M12 = vdm^T vdm (used by invmass)
Point sets are generated here: tensor Gauss(-Lobatto)-Legendre points on the
quadrilateral in x-fastest order (the order of the reference's tabulated
quad rules), Gauss-Legendre flux points on each face in ascending face
coordinate, and the Williams-Shunn-Jameson rules (J. Comput. Appl. Math. 266,
2014) on the triangle.
"""
import itertools as it
import re

import numpy as np

from dgfs_hip.polys import get_polybasis
from dgfs_hip.quadratures import zwgj, zwglj

# Williams-Shunn-Jameson symmetric triangle rules (points on [-1,1]^2 reference triangle)
_WS_TRI = {
    3: [(-0.666666666666667, 0.333333333333333), (0.333333333333333, -0.666666666666667),
        (-0.666666666666667, -0.666666666666667)],
    6: [(-0.81684757298044, 0.63369514596088), (0.63369514596088, -0.81684757298044),
        (-0.81684757298044, -0.81684757298044), (-0.108103018168072, -0.783793963663856),
        (-0.783793963663856, -0.108103018168072), (-0.108103018168072, -0.108103018168072)],
}


def line_rule(rule, npts):
    if rule == "gauss-legendre":
        return zwgj(npts, 0.0, 0.0)
    if rule == "gauss-legendre-lobatto":
        return zwglj(npts, 0.0, 0.0)
    raise ValueError(f"unknown line rule {rule!r}")


def _proj_pts(proj, pts):
    pts = np.atleast_2d(np.asarray(pts, dtype=float).T)
    return np.vstack(np.broadcast_arrays(*proj(*pts))).T


class BaseShape:
    name = None
    ndims = -1

    def __init__(self, order, upts_rule="gauss-legendre", fpts_rule="gauss-legendre"):
        self.order = order
        self.upts_rule = upts_rule
        self.fpts_rule = fpts_rule
        self.upts = self._upts()
        self.nupts = len(self.upts)
        self.ubasis = get_polybasis(self.name, order + 1, self.upts)
        # faces
        ppts, pnorms, nfacefpts = [], [], []
        z, _ = line_rule(fpts_rule, order + 1)
        self.face_pts1d = z
        for kind, proj, norm in self.faces:
            fp = _proj_pts(proj, z)
            ppts.append(fp)
            pnorms.append(np.tile(norm, (len(fp), 1)))
            nfacefpts.append(len(fp))
        self.fpts = np.vstack(ppts)
        self.norm_fpts = np.vstack(pnorms).astype(float)
        self.nfacefpts = nfacefpts
        self.nfpts = len(self.fpts)
        nf = np.cumsum([0] + nfacefpts)
        self.facefpts = [list(range(nf[i], nf[i + 1])) for i in range(len(nfacefpts))]

    # operator matrices ----------------------------------------------------
    @property
    def m0(self):
        return self.ubasis.nodal_basis_at(self.fpts)

    @property
    def m1(self):
        m = np.rollaxis(self.ubasis.jac_nodal_basis_at(self.upts), 2)
        return m.reshape(self.nupts, -1)

    @property
    def m2(self):
        m = self.norm_fpts[..., None] * self.m0[:, None, :]
        return m.reshape(self.nfpts, -1)

    @property
    def gbasis_coeffs(self):
        coeffs = []
        zq, wq = line_rule("gauss-legendre-lobatto", self.order + 1)
        fb = get_polybasis("line", self.order + 1, self.face_pts1d)
        for kind, proj, norm in self.faces:
            L = fb.nodal_basis_at(zq)                       # [q][k]
            M = self.ubasis.ortho_basis_at(_proj_pts(proj, zq))  # [j][q]
            coeffs.append(np.einsum("i,ik,ji->kj", wq, L, M))
        return np.vstack(coeffs)

    @property
    def m3(self):
        return (self.gbasis_coeffs @ self.ubasis.ortho_basis_at(self.upts)).T

    @property
    def m12(self):
        v = self.ubasis.vdm
        return (v.T @ v).reshape(self.nupts, -1)

    def opmat(self, expr):
        """Evaluate an operator expression such as 'M1 - M3*M2' (shapes.py:83-91)."""
        if not re.match(r"[M0-9\-+*() ]+$", expr):
            raise ValueError("Invalid operator matrix expression")
        mats = {m: np.asarray(getattr(self, m.lower())) for m in re.findall(r"M\d+", expr)}
        py = re.sub(r"(M\d+)\s*\*\s*(M\d+)", r"(\1 @ \2)", expr)
        return np.asarray(eval(py, {"__builtins__": None}, mats))

    # geometry: linear shape functions ------------------------------------
    def shape_basis(self):
        return get_polybasis(self.name, 2, self.std_ele_lin())


class QuadShape(BaseShape):
    name = "quad"
    ndims = 2
    faces = [
        ("line", lambda s: (s, -1), (0, -1)),
        ("line", lambda s: (1, s), (1, 0)),
        ("line", lambda s: (s, 1), (0, 1)),
        ("line", lambda s: (-1, s), (-1, 0)),
    ]

    def _upts(self):
        z, _ = line_rule(self.upts_rule, self.order + 1)
        return np.array([(z[i], z[j]) for j in range(len(z)) for i in range(len(z))])

    @staticmethod
    def std_ele_lin():
        pts1d = np.linspace(-1, 1, 2)
        return np.array([p[::-1] for p in it.product(pts1d, repeat=2)])


class TriShape(BaseShape):
    name = "tri"
    ndims = 2
    faces = [
        ("line", lambda s: (s, -1), (0, -1)),
        ("line", lambda s: (-s, s), (1, 1)),
        ("line", lambda s: (-1, s), (-1, 0)),
    ]

    def _upts(self):
        n = (self.order + 1) * (self.order + 2) // 2
        if self.upts_rule != "williams-shunn" or n not in _WS_TRI:
            raise ValueError(f"triangle solution points: williams-shunn with n in {sorted(_WS_TRI)}")
        return np.array(_WS_TRI[n])

    @staticmethod
    def std_ele_lin():
        return np.array([(-1.0, -1.0), (1.0, -1.0), (-1.0, 1.0)])


def get_shape(name, order, upts_rule=None, fpts_rule="gauss-legendre"):
    if name == "quad":
        return QuadShape(order, upts_rule or "gauss-legendre", fpts_rule)
    if name == "tri":
        return TriShape(order, upts_rule or "williams-shunn", fpts_rule)
    raise ValueError(f"unsupported element type {name!r}")
