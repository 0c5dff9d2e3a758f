"""Symmetric (antipodal) spherical designs on S^2.

Mirrors the selection rule of frfs/sphericaldesign/__init__.py:59-107: a rule
is requested by number of points (2M) and the velocity mesh keeps the first M
rows as the half-sphere set (frfs/solvers/dgfs/velocitymesh.py:128-135), so
row M+i must be the antipode of row i.

Built-in sets (closed forms, no data files):
  2 points   +-e_z                               (reference table ss001-m00002)
  6 points   the octahedron, rows e_z, e_x, e_y   (reference table ss003-m00006)
  12 points  the icosahedron: pole plus five points at z = 1/sqrt(5), azimuths
             0, -72, 144, -144, 72 degrees        (reference table ss005-m00012)
  24 points  every permutation of (0, +-a, +-b), a^4 + b^4 = 3/5 -- two
             icosahedra, an antipodal 5-design.  The reference ships no 24-point
             set; this one is chosen because it is exact to the same degree as
             the 12-point rule and closed-form.
An inline rule (a string of "x y z" lines, as get_sphquadrule accepts) may be
passed instead of a name.
"""
import math
import re

import numpy as np


class SphericalRule:
    def __init__(self, pts):
        self.pts = np.asarray(pts, dtype=float)
        self.wts = np.full(len(self.pts), 4.0 * math.pi / len(self.pts))


def _half(M):
    if M == 1:
        return [[0.0, 0.0, 1.0]]
    if M == 3:
        return [[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    if M == 6:
        zc, rc = 1.0 / math.sqrt(5.0), 2.0 / math.sqrt(5.0)
        out = [[0.0, 0.0, 1.0]]
        for deg in (0.0, -72.0, 144.0, -144.0, 72.0):
            t = math.radians(deg)
            out.append([rc * math.cos(t), rc * math.sin(t), zc])
        return out
    if M == 12:
        s = 0.5 * (1.0 + 1.0 / math.sqrt(5.0))
        a, b = math.sqrt(s), math.sqrt(1.0 - s)
        pts = []
        for u, v in ((a, b), (b, a)):
            for sv in (1.0, -1.0):
                pts += [[0.0, u, sv * v], [u, 0.0, sv * v], [u, sv * v, 0.0]]
        pts = np.array(pts)
        order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0]))[::-1]
        return pts[order].tolist()
    raise ValueError(f"No suitable spherical quadrature rule found ({2 * M} points)")


def get_sphquadrule(sstype, rule=None, npts=None, qdeg=None, flags=None):
    if sstype != "symmetric":
        raise ValueError(f"unknown spherical design family {sstype!r}")
    if rule and not re.match(r"[a-zA-Z0-9\-~+]+$", rule):
        rows = [list(map(float, ln.replace(",", " ").split()[:3]))
                for ln in rule.strip().splitlines() if ln.strip()]
        r = SphericalRule(rows)
        if npts and npts != len(r.pts):
            raise ValueError("Invalid number of points in provided rule")
        return r
    if not npts or npts % 2:
        raise ValueError("symmetric rules need an even number of points")
    half = np.array(_half(npts // 2))
    return SphericalRule(np.vstack([half, -half]))
