"""Reference elements and the FR operator matrices of the DGFS advection.

The operators and their layouts are the reference's (frfs/shapes.py:93-118):
  M0  (nfpts, nupts)          solution -> flux points: l_j(x_f)
  M1  (nupts, ndims*nupts)    reference gradient at the solution points,
                              column d*nupts + j = d l_j / d xi_d (x_u)
  M2  (nfpts, ndims*nupts)    reference face normal (unnormalised) times M0
  M3  (nupts, nfpts)          correction lift of the face fluxes
and opmat() evaluates expressions of them such as 'M1 - M3*M2'.

The lift M3 is the DG lift of shapes.py:243-289: column (face f, point k)
is the L2 projection onto the solution space of the face functional
q -> int_f l_k q ds, with the face integral taken by the (p+1)-point
Gauss-Lobatto rule.  With an orthonormal basis phi_b of the space,
M3[u, (f, k)] = sum_b phi_b(x_u) sum_q w_q l_k(s_q) phi_b(x_f(s_q)).

Point sets (conventions the meshes and the reference's files rely on):
* quadrilateral: tensor Gauss-Legendre (or -Lobatto) points, x fastest;
* triangle: the Williams-Shunn-Jameson points (J. Comput. Appl. Math. 266,
  2014), the tables of frfs/quadrules/tri/williams-shunn-n{3,6,10}-*.txt;
* faces: Gauss-Legendre points in ascending face coordinate, faces in the
  reference's order (quad: bottom, right, top, left; triangle: (s, -1),
  (-s, s), (-1, s)), with the reference's unnormalised normals.
"""
import re

import numpy as np

from dgfs_hip.polys import PolySpace, NodalBasis, nodal_basis
from dgfs_hip.quadratures import zwgj, zwglj

_WS_TRI = {
    3: [(-0.666666666666667, 0.333333333333333), (0.333333333333333, -0.666666666666667),
        (-0.666666666666667, -0.666666666666667)],
    6: [(-0.81684757298044, 0.63369514596088), (0.63369514596088, -0.81684757298044),
        (-0.81684757298044, -0.81684757298044), (-0.108103018168072, -0.783793963663856),
        (-0.783793963663856, -0.108103018168072), (-0.108103018168072, -0.108103018168072)],
    10: [(-0.333333333333333, -0.333333333333333), (-0.888871894660414, 0.777743789320828),
         (0.777743789320828, -0.888871894660414), (-0.888871894660414, -0.888871894660414),
         (0.268421495491446, -0.408932576528214), (-0.859488918963232, -0.408932576528214),
         (-0.408932576528214, -0.859488918963232), (0.268421495491446, -0.859488918963232),
         (-0.859488918963232, 0.268421495491446), (-0.408932576528214, 0.268421495491446)],
}


def line_rule(rule, npts):
    if rule == "gauss-legendre":
        return zwgj(npts, 0.0, 0.0)
    if rule == "gauss-legendre-lobatto":
        return zwglj(npts, 0.0, 0.0)
    raise ValueError(f"unknown line rule {rule!r}")


def _on_face(proj, s):
    """Reference-element points of the face parametrisation at s."""
    s = np.asarray(s, dtype=float)
    return np.stack([np.broadcast_to(c, s.shape).astype(float) for c in proj(s)], axis=1)


# -- operator expressions ----------------------------------------------------
_TOKEN = re.compile(r"\s*(M\d+|[-+*()])")


def eval_opmat(expr, lookup):
    """Value of an operator expression over matrices M<n>: '+', '-' (binary
    and unary), '*' (matrix product) and parentheses.  lookup(name) returns
    the matrix of a name."""
    if not re.fullmatch(r"[M0-9\-+*() ]+", expr):
        raise ValueError("Invalid operator matrix expression")
    toks, pos = [], 0
    while pos < len(expr):
        m = _TOKEN.match(expr, pos)
        if not m:
            if expr[pos:].strip():
                raise ValueError(f"Invalid operator matrix expression at {expr[pos:]!r}")
            break
        toks.append(m.group(1))
        pos = m.end()
    toks.append(None)
    at = [0]

    def peek():
        return toks[at[0]]

    def take():
        t = toks[at[0]]
        at[0] = min(at[0] + 1, len(toks) - 1)
        return t

    def factor():
        t = take()
        if t == "-":
            return -factor()
        if t == "(":
            v = sum_()
            if take() != ")":
                raise ValueError("unbalanced parentheses in operator expression")
            return v
        if t is not None and t.startswith("M"):
            return np.asarray(lookup(t), dtype=float)
        raise ValueError(f"unexpected {t!r} in operator expression")

    def product():
        v = factor()
        while peek() == "*":
            take()
            v = v @ factor()
        return v

    def sum_():
        v = product()
        while peek() in ("+", "-"):
            v = v + product() if take() == "+" else v - product()
        return v

    out = sum_()
    if peek() is not None:
        raise ValueError(f"trailing {peek()!r} in operator expression")
    return out


class BaseShape:
    name = None
    ndims = -1
    faces = ()

    def __init__(self, order, upts_rule="gauss-legendre", fpts_rule="gauss-legendre"):
        self.order = order
        self.upts_rule = upts_rule
        self.fpts_rule = fpts_rule
        self.upts = self._upts()
        self.nupts = len(self.upts)
        self.space = PolySpace(self.name, order)
        self.ubasis = NodalBasis(self.space, self.upts)
        self.face_pts1d, _ = line_rule(fpts_rule, order + 1)
        pts, nrm, counts = [], [], []
        for _, proj, norm in self.faces:
            fp = _on_face(proj, self.face_pts1d)
            pts.append(fp)
            nrm.append(np.tile(np.asarray(norm, dtype=float), (len(fp), 1)))
            counts.append(len(fp))
        self.fpts = np.vstack(pts)
        self.norm_fpts = np.vstack(nrm)
        self.nfacefpts = counts
        self.nfpts = len(self.fpts)
        edges = np.cumsum([0] + counts)
        self.facefpts = [list(range(edges[i], edges[i + 1])) for i in range(len(counts))]
        self._ops = {}

    # operator matrices ----------------------------------------------------
    def _op(self, name, build):
        if name not in self._ops:
            self._ops[name] = build()
        return self._ops[name]

    @property
    def m0(self):
        return self._op("m0", lambda: self.ubasis.at(self.fpts))

    @property
    def m1(self):
        def build():
            g = self.ubasis.grad_at(self.upts)            # (ndims, nupts_j, nupts_u)
            return np.concatenate([g[d].T for d in range(self.ndims)], axis=1)
        return self._op("m1", build)

    @property
    def m2(self):
        def build():
            return np.concatenate([self.norm_fpts[:, d:d + 1] * self.m0 for d in range(self.ndims)], axis=1)
        return self._op("m2", build)

    @property
    def m3(self):
        def build():
            zq, wq = line_rule("gauss-legendre-lobatto", self.order + 1)
            lk = nodal_basis("line", self.order, self.face_pts1d).at(zq)   # (nq, nk)
            phi_u = self.space.ortho_at(self.upts)                          # (nb, nupts)
            blocks = []
            for _, proj, _ in self.faces:
                phi_q = self.space.ortho_at(_on_face(proj, zq))             # (nb, nq)
                blocks.append(phi_u.T @ (phi_q @ (wq[:, None] * lk)))       # (nupts, nk)
            return np.hstack(blocks)
        return self._op("m3", build)

    def opmat(self, expr):
        """Operator expression such as 'M1 - M3*M2' (the syntax of shapes.py:83-91)."""
        return eval_opmat(expr, lambda m: getattr(self, m.lower()))

    # geometry: the linear shape functions --------------------------------
    def shape_basis(self):
        return nodal_basis(self.name, 1, self.std_ele_lin())


class QuadShape(BaseShape):
    name = "quad"
    ndims = 2
    faces = (
        ("line", lambda s: (s, -1.0), (0, -1)),
        ("line", lambda s: (1.0, s), (1, 0)),
        ("line", lambda s: (s, 1.0), (0, 1)),
        ("line", lambda s: (-1.0, s), (-1, 0)),
    )

    def _upts(self):
        z, _ = line_rule(self.upts_rule, self.order + 1)
        Y, X = np.meshgrid(z, z, indexing="ij")
        return np.stack([X.ravel(), Y.ravel()], axis=1)

    @staticmethod
    def std_ele_lin():
        return np.array([(-1.0, -1.0), (1.0, -1.0), (-1.0, 1.0), (1.0, 1.0)])


class TriShape(BaseShape):
    name = "tri"
    ndims = 2
    faces = (
        ("line", lambda s: (s, -1.0), (0, -1)),
        ("line", lambda s: (-s, s), (1, 1)),
        ("line", lambda s: (-1.0, s), (-1, 0)),
    )

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
