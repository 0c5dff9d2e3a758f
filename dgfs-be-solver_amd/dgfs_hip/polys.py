"""Polynomial spaces and nodal bases on the reference line, quadrilateral and
triangle.

The flux-reconstruction operators need two things of a solution space
(the role of frfs/polys.py in the reference):
* a nodal basis on the solution points (values and gradients anywhere), for
  the interpolation and differentiation operators;
* an L2-orthonormal basis of the same space on the reference element, for
  the DG lift of the face corrections (a projection, so any orthonormal basis
  gives the same operator).

Construction used here: the space is spanned by products of Legendre
polynomials P_i(x) P_j(y) (i, j <= p on the square, i + j <= p on the
triangle, i <= p on the line).  Their Gram matrix under an exact quadrature
rule of the reference element is factorised G = R^T R (Cholesky), and
phi = P R^-1 is orthonormal.  The nodal basis of a point set is the inverse
of the (square) matrix of any basis at the points.  Legendre values and
derivatives come from numpy.polynomial.legendre.

Reference elements: line [-1, 1]; quadrilateral [-1, 1]^2; triangle with
vertices (-1, -1), (1, -1), (-1, 1) (area 2).
"""
import numpy as np
from numpy.polynomial import legendre as _leg

_NDIMS = {"line": 1, "quad": 2, "tri": 2}


def _legendre(x, deg):
    """Values and first derivatives of P_0..P_deg at x: two (npts, deg+1)."""
    x = np.asarray(x, dtype=float).reshape(-1)
    val = _leg.legvander(x, deg)
    der = np.zeros_like(val)
    for i in range(1, deg + 1):
        c = np.zeros(i + 1)
        c[i] = 1.0
        der[:, i] = _leg.legval(x, _leg.legder(c))
    return val, der


def element_rule(shape, n):
    """Gauss rule with n points per direction on a reference element: exact
    for degree 2n - 1 per direction on the line and square, and for total
    degree 2n - 2 on the triangle (collapsed coordinates)."""
    z, w = _leg.leggauss(n)
    if shape == "line":
        return z[:, None], w
    A, B = np.meshgrid(z, z, indexing="ij")
    W = np.outer(w, w)
    if shape == "quad":
        return np.stack([A.ravel(), B.ravel()], axis=1), W.ravel()
    if shape == "tri":
        xi = (1.0 + A) * (1.0 - B) / 2.0 - 1.0
        return np.stack([xi.ravel(), B.ravel()], axis=1), (W * (1.0 - B) / 2.0).ravel()
    raise ValueError(f"no reference element {shape!r}")


class PolySpace:
    """P_p on the line or triangle, Q_p on the square, with an orthonormal
    basis phi_b (b = 0 .. size-1)."""

    def __init__(self, shape, degree):
        if shape not in _NDIMS:
            raise ValueError(f"no polynomial space on {shape!r}")
        self.shape, self.degree, self.ndims = shape, int(degree), _NDIMS[shape]
        p = self.degree
        if shape == "line":
            self.exps = [(i,) for i in range(p + 1)]
        elif shape == "quad":
            self.exps = [(i, j) for i in range(p + 1) for j in range(p + 1)]
        else:
            self.exps = [(i, j) for i in range(p + 1) for j in range(p + 1 - i)]
        self.size = len(self.exps)
        qp, qw = element_rule(shape, p + 2)
        raw = self._raw(qp)[0]
        gram = (raw * qw[:, None]).T @ raw
        R = np.linalg.cholesky(gram).T
        self._coef = np.linalg.inv(R)             # phi = raw @ coef

    def _raw(self, pts):
        """Legendre products and their gradients: (npts, size) and
        (ndims, npts, size)."""
        pts = np.asarray(pts, dtype=float).reshape(-1, self.ndims)
        axes = [_legendre(pts[:, d], self.degree) for d in range(self.ndims)]
        val = np.ones((len(pts), self.size))
        grad = np.ones((self.ndims, len(pts), self.size))
        for b, e in enumerate(self.exps):
            for d, i in enumerate(e):
                v, dv = axes[d][0][:, i], axes[d][1][:, i]
                val[:, b] *= v
                for g in range(self.ndims):
                    grad[g, :, b] *= dv if g == d else v
        return val, grad

    def ortho_at(self, pts):
        """(size, npts): phi_b at the points."""
        return (self._raw(pts)[0] @ self._coef).T

    def grad_ortho_at(self, pts):
        """(ndims, size, npts): d phi_b / d x_d at the points."""
        g = self._raw(pts)[1]
        return np.stack([(g[d] @ self._coef).T for d in range(self.ndims)])


class NodalBasis:
    """Lagrange basis of a PolySpace on a unisolvent point set: l_j(x_i) = delta_ij."""

    def __init__(self, space, nodes):
        self.space = space
        self.nodes = np.asarray(nodes, dtype=float).reshape(-1, space.ndims)
        if len(self.nodes) != space.size:
            raise ValueError(f"{len(self.nodes)} nodes for a space of dimension {space.size}")
        V = space.ortho_at(self.nodes)            # V[b, i] = phi_b(x_i)
        self._C = np.linalg.inv(V)                # l_j = sum_b C[j, b] phi_b, as C V = I

    def at(self, pts):
        """(npts, nnodes): l_j at every point."""
        return (self._C @ self.space.ortho_at(pts)).T

    def grad_at(self, pts):
        """(ndims, nnodes, npts): gradient of l_j along reference axis d."""
        G = self.space.grad_ortho_at(pts)
        return np.stack([self._C @ G[d] for d in range(self.space.ndims)])

    def ortho_at(self, pts):
        return self.space.ortho_at(pts)


def nodal_basis(shape, degree, nodes):
    return NodalBasis(PolySpace(shape, degree), nodes)
