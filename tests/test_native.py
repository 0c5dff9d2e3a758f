"""Solution files (writers/native.py, readers/native.py layout): container
round trips, numbering, restart interpolation onto another order, mesh
identity.  CPU only."""
import os
import types

import numpy as np
import pytest

from dgfs_hip.native import NativeReader, NativeWriter, dist_fields, restart_stats, soln_for_system, write_native


def _h5py_present():
    try:
        import h5py  # noqa: F401
        return True
    except ImportError:
        return False


CONTAINERS = ["npz"] + (["hdf5"] if _h5py_present() else [])


@pytest.mark.parametrize("container", CONTAINERS)
def test_roundtrip(tmp_path, container):
    rng = np.random.default_rng(0)
    a = rng.random((9, 5, 7))
    b = rng.random((4, 5, 3))
    meta = {"config": "[solver]\norder = 2\n", "stats": "[data]\nfields = f_0\n", "mesh_uuid": "abc"}
    p = write_native(str(tmp_path / "s.frfss"), {"soln_quad_p0": a, "soln_tri_p1": b}, meta, container)
    r = NativeReader(p)
    assert set(r) == {"soln_quad_p0", "soln_tri_p1", "config", "stats", "mesh_uuid"}
    assert np.array_equal(r["soln_quad_p0"], a) and r["soln_quad_p0"].dtype == np.float64
    assert r["config"] == meta["config"] and r["mesh_uuid"] == "abc"
    assert list(r.array_info("soln")) == ["soln_quad_p0", "soln_tri_p1"]
    assert r.partition_info("soln") == {"quad": [7, 0], "tri": [0, 3]}
    with pytest.raises(KeyError):
        r["nope"]
    r.close()
    assert not os.path.exists(p + ".tmp")


def test_hdf5_needs_h5py(tmp_path):
    if _h5py_present():
        pytest.skip("h5py is importable here")
    with pytest.raises(RuntimeError, match="h5py"):
        write_native(str(tmp_path / "x.frfss"), {"a": np.zeros(2)}, {}, "hdf5")
    p = tmp_path / "fake.frfss"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(RuntimeError, match="h5py"):
        NativeReader(str(p))


def test_reader_refuses_other_files_and_pickles(tmp_path):
    p = tmp_path / "x.frfss"
    p.write_text("not a solution")
    with pytest.raises(ValueError):
        NativeReader(str(p))
    q = tmp_path / "obj.frfss"
    with open(q, "wb") as fh:
        np.savez(fh, obj=np.array([{"a": 1}], dtype=object))
    r = NativeReader(str(q))
    with pytest.raises(ValueError):
        r["obj"]


def test_writer_numbering_and_restart(tmp_path):
    w = NativeWriter(str(tmp_path), "run-{n:03d}", prefix="soln")
    paths = [w.write({"quad": np.full((1, 1, 1), float(i))}, {"stats": ""}, 0.1 * i) for i in range(3)]
    assert [os.path.basename(p) for p in paths] == ["run-000.frfss", "run-001.frfss", "run-002.frfss"]
    assert NativeReader(paths[2])["soln_quad_p0"][0, 0, 0] == 2.0
    w2 = NativeWriter(str(tmp_path), "run-{n:03d}", prefix="soln", isrestart=True)
    assert w2.nout == 3
    wt = NativeWriter(str(tmp_path), "t-{t:.2f}", prefix="soln")
    assert os.path.basename(wt.write({"quad": np.zeros((1, 1, 1))}, {}, 0.25)) == "t-0.25.frfss"


def test_dist_fields():
    assert dist_fields(3) == "f_0,f_1,f_2"
    assert dist_fields(2, 2) == "f_0:1,f_1:1,f_0:2,f_1:2"


def _file_with(tmp_path, soln, order, fields, etype="quad", tcurr=0.5, nsteps=7):
    from dgfs_hip.inifile import Inifile
    cfg = Inifile()
    cfg.set("solver", "order", order)
    st = Inifile()
    st.set("data", "fields", fields)
    st.set("solver-time-integrator", "tcurr", repr(tcurr))
    st.set("solver-time-integrator", "nsteps", nsteps)
    p = write_native(str(tmp_path / "r.frfss"), {f"soln_{etype}_p0": soln},
                     {"config": cfg.tostr(), "stats": st.tostr(), "mesh_uuid": "m"})
    return NativeReader(p)


@pytest.mark.parametrize("etype,p0,p1", [("quad", 1, 2), ("quad", 2, 3), ("tri", 1, 2), ("quad", 2, 2)])
def test_restart_interpolates_polynomials_exactly(tmp_path, etype, p0, p1):
    """A field of degree p0 stored at order p0 reads back at order p1 >= p0
    as the same polynomial at the new points (solvers/dgfs/system.py:98-108)."""
    from dgfs_hip.shapes import get_shape
    rule = "gauss-legendre" if etype == "quad" else "williams-shunn"
    s0, s1 = get_shape(etype, p0, rule), get_shape(etype, p1, rule)

    def poly(x, k):
        return (1 + k) + 0.5 * x[:, 0] - 0.25 * k * x[:, 1] + 0.1 * x[:, 0] * x[:, 1] ** (p0 - 1)

    neles, nv = 3, 2
    soln = np.stack([np.stack([poly(s0.upts, e + v) for v in range(nv)], 1) for e in range(neles)], 2)
    r = _file_with(tmp_path, soln, p0, dist_fields(nv), etype)
    ele = types.SimpleNamespace(nupts=s1.nupts, neles=neles, basis=s1)
    got = soln_for_system(r, None, ele, dist_fields(nv), etype)
    want = np.stack([np.stack([poly(s1.upts, e + v) for v in range(nv)], 1) for e in range(neles)], 2)
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-13
    if p0 == p1:
        assert got is not None and np.array_equal(got, soln)
    assert restart_stats(r) == (0.5, 7)


def test_restart_checks_fields_and_size(tmp_path):
    from dgfs_hip.shapes import get_shape
    s = get_shape("quad", 1, "gauss-legendre")
    r = _file_with(tmp_path, np.zeros((4, 2, 3)), 1, dist_fields(2))
    ele = types.SimpleNamespace(nupts=4, neles=3, basis=s)
    with pytest.raises(RuntimeError, match="system"):
        soln_for_system(r, None, ele, dist_fields(3), "quad")
    with pytest.raises(RuntimeError, match="mesh"):
        soln_for_system(r, None, types.SimpleNamespace(nupts=4, neles=5, basis=s), dist_fields(2), "quad")


def test_mesh_uuid():
    from dgfs_hip.mesh import rect_quads, rect_tris
    a, b = rect_quads(3, 2), rect_quads(3, 2)
    assert a.uuid == b.uuid
    assert rect_quads(3, 2, periodic_x=True).uuid != a.uuid
    assert rect_quads(3, 2, x1=1.1).uuid != a.uuid
    assert rect_tris(3, 2).uuid != a.uuid
