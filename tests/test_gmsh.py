"""Gmsh (MSH 2.2 ASCII) import: a mesh written from the structured builders
reads back as the same elements, interfaces, periodic pairs and boundaries
(frfs/readers/gmsh.py and readers/base.py connectivity).  CPU only."""
import numpy as np
import pytest

from dgfs_hip.gmsh import read_gmsh
from dgfs_hip.mesh import rect_quads, rect_tris

_GMSH_ORDER = {"quad": [0, 1, 3, 2], "tri": [0, 1, 2]}   # inverse of the node map (self-inverse)
_GMSH_TYPE = {"quad": 3, "tri": 2}
_FACE_VERTS = {"quad": [(0, 1), (1, 2), (2, 3), (3, 0)], "tri": [(0, 1), (1, 2), (2, 0)]}


def write_msh(mesh, names, rename=None, shuffle_seed=None):
    """MSH 2.2 text of a structured Mesh.  names: boundary names to emit as
    physical lines; rename maps a boundary to another physical name."""
    rename = rename or {}
    et = mesh.eletype
    ids, pts = {}, []
    gverts = []
    for e in range(mesh.neles):
        gv = []
        for k in _GMSH_ORDER[et]:
            p = tuple(float(x) for x in np.round(mesh.nodes[k, e], 12))
            if p not in ids:
                ids[p] = len(ids) + 1
                pts.append(p)
            gv.append(ids[p])
        gverts.append(gv)
    phys = {"fluid": (2, 1)}
    for i, n in enumerate(names):
        phys[rename.get(n, n)] = (1, i + 2)
    out = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(phys))]
    out += [f'{d} {t} "{n}"' for n, (d, t) in phys.items()]
    out += ["$EndPhysicalNames", "$Nodes", str(len(pts))]
    out += [f"{i + 1} {p[0]!r} {p[1]!r} 0" for i, p in enumerate(pts)]
    out += ["$EndNodes"]
    eles = []
    for n in names:
        tag = phys[rename.get(n, n)][1]
        for (e, f) in mesh.boundaries[n]:
            a, b = _FACE_VERTS[et][f]
            eles.append(f"1 2 {tag} {tag} {gverts[e][a]} {gverts[e][b]}")
    for e in range(mesh.neles):
        eles.append(f"{_GMSH_TYPE[et]} 2 1 1 " + " ".join(map(str, gverts[e])))
    if shuffle_seed is not None:
        np.random.default_rng(shuffle_seed).shuffle(eles[: -mesh.neles])
    out += ["$Elements", str(len(eles))]
    out += [f"{i + 1} {l}" for i, l in enumerate(eles)]
    out += ["$EndElements", "$Comments", "skipped", "$EndComments"]
    return out


def _iface_set(inters):
    s = {}
    for (eL, fL, eR, fR, shift) in inters:
        a, b = (eL, fL), (eR, fR)
        sh = np.asarray(shift, float)
        if b < a:
            a, b, sh = b, a, -sh
        s[(a, b)] = tuple(np.round(sh, 12) + 0.0)
    return s


def _same_mesh(m, ref):
    assert m.eletype == ref.eletype
    assert np.allclose(m.nodes, ref.nodes, rtol=0, atol=1e-12)
    assert _iface_set(m.interfaces) == _iface_set(ref.interfaces)
    assert {k: sorted(v) for k, v in m.boundaries.items()} == {k: sorted(v) for k, v in ref.boundaries.items()}


@pytest.mark.parametrize("builder", [rect_quads, rect_tris])
def test_gmsh_roundtrip_walls(builder):
    ref = builder(5, 3, x0=-1.0, x1=2.0, y0=0.5, y1=1.5)
    m = read_gmsh(write_msh(ref, ["bottom", "top", "left", "right"], shuffle_seed=3))
    _same_mesh(m, ref)


@pytest.mark.parametrize("builder", [rect_quads, rect_tris])
def test_gmsh_periodic_pairs(builder):
    walls = builder(4, 3, x1=2.0)
    ref = builder(4, 3, x1=2.0, periodic_x=True)
    txt = write_msh(walls, ["bottom", "top", "left", "right"],
                    rename={"left": "periodic_0_r", "right": "periodic_0_l"})
    m = read_gmsh(txt)
    _same_mesh(m, ref)


def test_gmsh_names_are_case_insensitive_and_files_read(tmp_path):
    ref = rect_quads(2, 2)
    txt = write_msh(ref, ["bottom", "top", "left", "right"], rename={"top": "Top"})
    p = tmp_path / "box.msh"
    p.write_text("\n".join(txt) + "\n")
    m = read_gmsh(str(p))
    _same_mesh(m, ref)


def test_gmsh_errors():
    ref = rect_quads(2, 2)
    good = write_msh(ref, ["bottom", "top", "left", "right"])
    with pytest.raises(ValueError, match="format"):
        read_gmsh([l.replace("2.2 0 8", "4.1 0 8") for l in good])
    with pytest.raises(ValueError, match="neither interior"):
        read_gmsh(write_msh(ref, ["bottom", "top", "left"]))
    with pytest.raises(ValueError, match="No fluid"):
        read_gmsh([l.replace('"fluid"', '"solid"') for l in good])
    with pytest.raises(ValueError, match="Unpaired periodic"):
        read_gmsh(write_msh(ref, ["bottom", "top", "left", "right"], rename={"left": "periodic_a_l"}))
    with pytest.raises(ValueError, match="Required section"):
        i, j = good.index("$Nodes"), good.index("$EndNodes")
        read_gmsh(good[:i] + good[j + 1:])


def test_annulus_mesh_is_consistent():
    """The cylinder mesh: counter-clockwise triangles, every face either
    paired once with a face of the same end points or on one boundary, the
    annulus area, and a round trip through the gmsh reader."""
    from dgfs_hip.mesh import annulus_tris
    m = annulus_tris(4, 12, r0=0.5, r1=3.0, growth=1.2)
    x = m.nodes
    area = 0.5 * ((x[1, :, 0] - x[0, :, 0]) * (x[2, :, 1] - x[0, :, 1])
                  - (x[2, :, 0] - x[0, :, 0]) * (x[1, :, 1] - x[0, :, 1]))
    assert np.all(area > 0)
    poly = 0.5 * 12 * np.sin(2 * np.pi / 12)
    assert abs(area.sum() - poly * (3.0 ** 2 - 0.5 ** 2)) < 1e-12
    seen = {}
    ends = lambda e, f: {tuple(np.round(x[a, e], 12)) for a in _FACE_VERTS["tri"][f]}
    for (eL, fL, eR, fR, sh) in m.interfaces:
        assert ends(eL, fL) == ends(eR, fR)
        for k in ((eL, fL), (eR, fR)):
            assert k not in seen
            seen[k] = 1
    for name, faces in m.boundaries.items():
        for k in faces:
            assert k not in seen
            seen[k] = 1
    assert len(seen) == 3 * m.neles
    r = lambda faces: [np.linalg.norm(x[_FACE_VERTS["tri"][f][0], e]) for e, f in faces]
    assert np.allclose(r(m.boundaries["cylinder"]), 0.5) and np.allclose(r(m.boundaries["inlet"]), 3.0)
    assert len(m.boundaries["inlet"]) + len(m.boundaries["outlet"]) == 12
    back = read_gmsh(write_msh(m, ["cylinder", "inlet", "outlet"]))
    _same_mesh(back, m)
