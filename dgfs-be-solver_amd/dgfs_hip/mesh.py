"""In-process structured meshes: rectangles of quadrilaterals or of
triangles (each quad split along its diagonal), with optional periodic
pairing of the x-faces.  Replaces mesh import (frfs/readers) for the
structured configurations; element nodes follow the reference's linear node
ordering (TensorProdShape.std_ele / TriShape.std_ele)."""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    eletype: str
    nodes: np.ndarray                      # (nspts, neles, ndims)
    interfaces: list                       # (eL, fL, eR, fR, shift(ndims))
    boundaries: dict = field(default_factory=dict)   # name -> [(e, f)]

    @property
    def neles(self):
        return self.nodes.shape[1]

    @property
    def uuid(self):
        """Identity of the mesh for solution files (the reference's
        mesh_uuid): a name-based UUID of the element type, nodes,
        interfaces and boundaries, so equal meshes share it."""
        import hashlib
        import uuid
        h = hashlib.sha256(self.eletype.encode())
        h.update(np.ascontiguousarray(self.nodes, dtype=float).tobytes())
        for (eL, fL, eR, fR, sh) in self.interfaces:
            h.update(np.array([eL, fL, eR, fR], dtype=np.int64).tobytes())
            h.update(np.asarray(sh, dtype=float).tobytes())
        for name in sorted(self.boundaries):
            h.update(name.encode())
            h.update(np.array(self.boundaries[name], dtype=np.int64).tobytes())
        return str(uuid.UUID(bytes=h.digest()[:16], version=5))


def rect_quads(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0, periodic_x=False, periodic_y=False):
    """Quad faces: 0 bottom, 1 right, 2 top, 3 left; element e = ix + nx*iy.
    Boundaries: 'bottom'/'top' unless periodic in y, 'left'/'right' unless
    periodic in x."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    neles = nx * ny
    nodes = np.empty((4, neles, 2))
    for iy in range(ny):
        for ix in range(nx):
            e = ix + nx * iy
            nodes[:, e] = [(xs[ix], ys[iy]), (xs[ix + 1], ys[iy]), (xs[ix], ys[iy + 1]), (xs[ix + 1], ys[iy + 1])]
    inters, bnd = [], {"bottom": [], "top": [], "left": [], "right": []}
    zero = np.zeros(2)
    for iy in range(ny):
        for ix in range(nx):
            e = ix + nx * iy
            if ix + 1 < nx:
                inters.append((e, 1, e + 1, 3, zero))
            elif periodic_x:
                inters.append((e, 1, nx * iy, 3, np.array([x1 - x0, 0.0])))
            else:
                bnd["right"].append((e, 1))
            if ix == 0 and not periodic_x:
                bnd["left"].append((e, 3))
            if iy + 1 < ny:
                inters.append((e, 2, e + nx, 0, zero))
            elif periodic_y:
                inters.append((e, 2, ix, 0, np.array([0.0, y1 - y0])))
            else:
                bnd["top"].append((e, 2))
            if iy == 0 and not periodic_y:
                bnd["bottom"].append((e, 0))
    if periodic_x:
        del bnd["left"], bnd["right"]
    if periodic_y:
        del bnd["bottom"], bnd["top"]
    return Mesh("quad", nodes, inters, bnd)


def rect_tris(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0, periodic_x=False):
    """Each quad (ix, iy) becomes a lower-right triangle (2*q) with nodes
    (x0,y0),(x1,y0),(x1,y1) and an upper-left one (2*q+1) with nodes
    (x1,y1),(x0,y1),(x0,y0).  Triangle faces follow TriShape: 0 between
    nodes 0-1, 1 between nodes 1-2, 2 between nodes 2-0."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    nq = nx * ny
    nodes = np.empty((3, 2 * nq, 2))
    for iy in range(ny):
        for ix in range(nx):
            q = ix + nx * iy
            a, b, c, d = (xs[ix], ys[iy]), (xs[ix + 1], ys[iy]), (xs[ix + 1], ys[iy + 1]), (xs[ix], ys[iy + 1])
            nodes[:, 2 * q] = [a, b, c]
            nodes[:, 2 * q + 1] = [c, d, a]
    inters, bnd = [], {"bottom": [], "top": [], "left": [], "right": []}
    zero = np.zeros(2)
    for iy in range(ny):
        for ix in range(nx):
            q = ix + nx * iy
            lo, up = 2 * q, 2 * q + 1
            inters.append((lo, 2, up, 2, zero))          # diagonal
            if ix + 1 < nx:
                inters.append((lo, 1, 2 * (q + 1) + 1, 1, zero))
            elif periodic_x:
                inters.append((lo, 1, 2 * (nx * iy) + 1, 1, np.array([x1 - x0, 0.0])))
            else:
                bnd["right"].append((lo, 1))
            if ix == 0 and not periodic_x:
                bnd["left"].append((up, 1))
            if iy + 1 < ny:
                inters.append((up, 0, 2 * (q + nx), 0, zero))
            else:
                bnd["top"].append((up, 0))
            if iy == 0:
                bnd["bottom"].append((lo, 0))
    if periodic_x:
        del bnd["left"], bnd["right"]
    return Mesh("tri", nodes, inters, bnd)


def annulus_tris(nr, ntheta, r0=0.5, r1=5.0, growth=1.1, flow=(1.0, 0.0)):
    """Triangles around a cylinder of radius r0 out to a circular far field of
    radius r1: nr rings with radial spacing growing by `growth`, ntheta
    sectors, each ring cell split along its diagonal into two triangles
    (counter-clockwise nodes, faces as in rect_tris).  Boundaries:
    'cylinder' (r = r0), and on r = r1 'inlet' where the outward normal
    faces the flow direction and 'outlet' elsewhere.  Sides are straight,
    so the cylinder is a regular ntheta-gon."""
    dr = np.array([growth ** i for i in range(nr)])
    rs = r0 + (r1 - r0) * np.concatenate([[0.0], np.cumsum(dr)]) / dr.sum()
    th = 2 * np.pi * np.arange(ntheta) / ntheta
    P = np.stack([rs[:, None] * np.cos(th)[None, :], rs[:, None] * np.sin(th)[None, :]], axis=-1)
    nq = nr * ntheta
    nodes = np.empty((3, 2 * nq, 2))
    for i in range(nr):
        for j in range(ntheta):
            q = j + ntheta * i
            jn = (j + 1) % ntheta
            a, b, c, d = P[i, j], P[i + 1, j], P[i + 1, jn], P[i, jn]
            nodes[:, 2 * q] = [a, b, c]
            nodes[:, 2 * q + 1] = [c, d, a]
    fd = np.asarray(flow, dtype=float)
    inters, bnd = [], {"cylinder": [], "inlet": [], "outlet": []}
    zero = np.zeros(2)
    # lo = [a, b, c]: face 0 the sector edge at theta_j, face 1 the ring edge at
    # r_{i+1}, face 2 the diagonal; up = [c, d, a]: face 0 the sector edge at
    # theta_{j+1}, face 1 the ring edge at r_i, face 2 the diagonal
    for i in range(nr):
        for j in range(ntheta):
            q = j + ntheta * i
            lo, up = 2 * q, 2 * q + 1
            inters.append((lo, 2, up, 2, zero))
            if i + 1 < nr:
                inters.append((lo, 1, 2 * (q + ntheta) + 1, 1, zero))
            else:
                mid = 0.5 * (P[nr, j] + P[nr, (j + 1) % ntheta])
                bnd["inlet" if mid @ fd < 0 else "outlet"].append((lo, 1))
            qn = (j + 1) % ntheta + ntheta * i
            inters.append((up, 0, 2 * qn, 0, zero))
            if i == 0:
                bnd["cylinder"].append((up, 1))
    return Mesh("tri", nodes, inters, bnd)
