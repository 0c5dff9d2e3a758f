"""Gmsh mesh import (MSH 2.2 ASCII) for the HIP path.

Follows frfs/readers/gmsh.py:31-204 and the connectivity assembly of
frfs/readers/base.py:80-200: physical names select the fluid elements
('fluid'), the boundaries (every other name, which becomes the
`soln-bcs-<name>` section) and periodic face pairs ('periodic_<id>_l' /
'periodic_<id>_r'); faces are paired through their sorted vertex ids, and
periodic faces in the order of their centroids.  Elements must be of one 2D
type (triangles or quadrangles; higher-order elements keep their vertices,
i.e. straight sides).  The result
is the Mesh of dgfs_hip.mesh, with the reference's linear node ordering
(gmsh quad nodes 0 1 2 3 -> tensor order 0 1 3 2) and face numbering
(face k between gmsh vertices k and k+1).
"""
import re
from collections import defaultdict

import numpy as np

from dgfs_hip.mesh import Mesh

# gmsh element type -> (shape, vertex count)
_ETYPES = {1: ("line", 2), 8: ("line", 2), 26: ("line", 2), 27: ("line", 2),
           2: ("tri", 3), 9: ("tri", 3), 21: ("tri", 3), 23: ("tri", 3),
           3: ("quad", 4), 10: ("quad", 4), 36: ("quad", 4), 37: ("quad", 4)}
# faces as gmsh vertex pairs (readers/gmsh.py:48-51)
_FACES = {"tri": [(0, 1), (1, 2), (2, 0)], "quad": [(0, 1), (1, 2), (2, 3), (3, 0)]}
# gmsh vertex order -> the reference's linear node order (std_ele)
_NODEMAP = {"tri": [0, 1, 2], "quad": [0, 1, 3, 2]}


def _sections(lines):
    it = iter(lines)
    for l in it:
        l = l.strip()
        if not l:
            continue
        if not l.startswith("$"):
            raise ValueError("Expected a mesh section")
        name = l[1:]
        body = []
        for b in it:
            b = b.strip()
            if b == "$End" + name:
                break
            body.append(b)
        else:
            raise ValueError("Expected $End" + name)
        yield name, body


def _fuzzysort(pts, idx, dim=0, tol=1e-6):
    """Indices idx ordered by pts[dim], values within tol counting as equal
    and ordered by the next dimension (the ordering of frfs/nputil.py:84-100
    that pairs periodic faces)."""
    srt = sorted(idx, key=lambda i: pts[dim][i])
    if dim + 1 >= len(pts) or len(srt) < 2:
        return srt
    out, i = [], 0
    while i < len(srt):
        j = i + 1
        while j < len(srt) and pts[dim][srt[j]] - pts[dim][srt[i]] < tol:
            j += 1
        grp = srt[i:j]
        out.extend(_fuzzysort(pts, grp, dim + 1, tol) if len(grp) > 1 else grp)
        i = j
    return out


def read_gmsh(path_or_lines, ndims=2):
    if isinstance(path_or_lines, str):
        with open(path_or_lines) as fh:
            lines = fh.read().splitlines()
    else:
        lines = list(path_or_lines)
    sect = dict(_sections(lines))
    for req in ("MeshFormat", "PhysicalNames", "Nodes", "Elements"):
        if req not in sect:
            raise ValueError(f"Required section {req} not found")
    ver, ftype, dsize = sect["MeshFormat"][0].split()
    if ver != "2.2" or ftype != "0" or dsize != "8":
        raise ValueError("Invalid mesh format: MSH 2.2 ASCII expected")

    fluid, bnames, periodic = None, {}, defaultdict(dict)
    for l in sect["PhysicalNames"][1:]:
        m = re.match(r'(\d+) (\d+) "((?:[^"\\]|\\.)*)"$', l)
        if not m:
            raise ValueError("Malformed physical entity")
        tag, name = int(m.group(2)), m.group(3).lower()
        if name == "fluid":
            fluid = tag
        elif name.startswith("periodic"):
            p = re.match(r"periodic[ _-]([a-z0-9]+)[ _-](l|r)$", name)
            if not p:
                raise ValueError("Invalid periodic boundary condition")
            periodic[p.group(1)][p.group(2)] = tag
        else:
            bnames[tag] = name
    if fluid is None:
        raise ValueError("No fluid elements in mesh")
    if any(len(v) != 2 for v in periodic.values()):
        raise ValueError("Unpaired periodic boundary in mesh")

    pts = {}
    for l in sect["Nodes"][1:]:
        v = l.split()
        pts[int(v[0])] = np.array([float(x) for x in v[1:1 + ndims]])

    eles, etype, lines_by_tag = [], None, defaultdict(list)
    for l in sect["Elements"][1:]:
        v = [int(x) for x in l.split()]
        et, ntags = v[1], v[2]
        if et not in _ETYPES:
            raise ValueError(f"Unsupported element type {et}")
        shape, nv = _ETYPES[et]
        tag, verts = v[3], v[3 + ntags:3 + ntags + nv]
        if shape == "line":
            lines_by_tag[tag].append(tuple(verts))
        elif tag == fluid:
            if etype not in (None, shape):
                raise ValueError("mixed element types are not supported")
            etype = shape
            eles.append(verts)
    if not eles:
        raise ValueError("No fluid elements in mesh")

    neles = len(eles)
    nodes = np.empty((len(_NODEMAP[etype]), neles, ndims))
    for e, verts in enumerate(eles):
        for k, g in enumerate(_NODEMAP[etype]):
            nodes[k, e] = pts[verts[g]]

    # faces through sorted vertex ids (readers/base.py:122-136)
    open_faces = {}
    inters = []
    zero = np.zeros(ndims)
    for e, verts in enumerate(eles):
        for f, (a, b) in enumerate(_FACES[etype]):
            key = tuple(sorted((verts[a], verts[b])))
            if key in open_faces:
                eL, fL = open_faces.pop(key)
                inters.append((eL, fL, e, f, zero))
            else:
                open_faces[key] = (e, f)

    def centroid(key):
        return 0.5 * (pts[key[0]] + pts[key[1]])

    # periodic pairs: both sides' faces ordered by their centroids, the
    # i-th left face paired with the i-th right face (readers/base.py:138-161)
    for pid, tags in periodic.items():
        lk = [tuple(sorted(v)) for v in lines_by_tag[tags["l"]]]
        rk = [tuple(sorted(v)) for v in lines_by_tag[tags["r"]]]
        if len(lk) != len(rk) or not lk:
            raise ValueError(f"periodic boundary {pid}: sides differ in face count")
        cl = np.array([centroid(k) for k in lk])
        cr = np.array([centroid(k) for k in rk])
        li = _fuzzysort(cl.T, range(len(lk)))
        ri = _fuzzysort(cr.T, range(len(rk)))
        for a, b in zip(li, ri):
            if lk[a] not in open_faces or rk[b] not in open_faces:
                raise ValueError(f"periodic boundary {pid}: face is not a free face of the fluid")
            eL, fL = open_faces.pop(lk[a])
            eR, fR = open_faces.pop(rk[b])
            # shift carries the right-hand face onto the left-hand one
            inters.append((eL, fL, eR, fR, cl[a] - cr[b]))

    boundaries = {}
    for tag, name in bnames.items():
        faces = []
        for v in lines_by_tag.get(tag, []):
            key = tuple(sorted(v))
            if key not in open_faces:
                raise ValueError(f"boundary {name}: face {key} is not a free face of the fluid")
            faces.append(open_faces.pop(key))
        if faces:
            boundaries[name] = faces
    if open_faces:
        raise ValueError(f"{len(open_faces)} faces are neither interior, periodic nor on a named boundary")
    return Mesh(etype, nodes, inters, boundaries)
