"""Gauss-Jacobi and Gauss-Lobatto-Jacobi rules by Golub-Welsch.

Mirrors frfs/quadratures.py:6-55 (zwgj, zwglj): nodes are the eigenvalues of
the symmetric Jacobi matrix of the monic Jacobi recurrence, weights come from
the first eigenvector components.  The Lobatto variant modifies the last
recurrence pair so that the degree-n polynomial vanishes at -1 and +1.
"""
import math

import numpy as np


def _recurrence(n, a, b):
    alpha = np.zeros(n)
    beta = np.zeros(n)
    for k in range(n):
        s = 2.0 * k + a + b
        alpha[k] = (b - a) / (a + b + 2.0) if k == 0 else (b * b - a * a) / (s * (s + 2.0))
    beta[0] = 2.0 ** (a + b + 1.0) * math.gamma(a + 1.0) * math.gamma(b + 1.0) / math.gamma(a + b + 2.0)
    for k in range(1, n):
        s = 2.0 * k + a + b
        beta[k] = 4.0 * k * (k + a) * (k + b) * (k + a + b) / (s * s * (s + 1.0) * (s - 1.0))
    return alpha, beta


def _eig(alpha, beta):
    off = np.sqrt(beta[1:])
    J = np.diag(alpha) + np.diag(off, 1) + np.diag(off, -1)
    z, V = np.linalg.eigh(J)
    order = np.argsort(z)
    return z[order], beta[0] * V[0, order] ** 2


def zwgj(n, a=0.0, b=0.0):
    """Gauss-Jacobi nodes and weights on [-1, 1]."""
    return _eig(*_recurrence(n, a, b))


def zwglj(n, a=0.0, b=0.0):
    """Gauss-Lobatto-Jacobi nodes and weights on [-1, 1] (n >= 2)."""
    if n < 2:
        raise ValueError("Gauss-Lobatto-Jacobi needs n >= 2")
    alpha, beta = _recurrence(n, a, b)

    def tail(x):
        pm2, pm1 = 1.0, x - alpha[0]
        for k in range(1, n - 1):
            pm2, pm1 = pm1, (x - alpha[k]) * pm1 - beta[k] * pm2
        return pm2, pm1

    lo, hi = tail(-1.0), tail(1.0)
    A = np.array([[lo[1], lo[0]], [hi[1], hi[0]]])
    al, be = np.linalg.solve(A, np.array([-lo[1], hi[1]]))
    alpha[n - 1], beta[n - 1] = al, be
    return _eig(alpha, beta)
