"""The C-ABI library loads and exports every symbol include/*.h declares.
CPU only: no compute call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dgfs_\w+|frfs_\w+)\s*\(", txt, flags=re.M))
    return syms


def test_library_exports_all_declared_symbols():
    from dgfs_hip import _lib
    path = _lib.lib_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    syms = declared_symbols()
    assert len(syms) >= 8
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_declared_symbols():
    from dgfs_hip import _lib
    from dgfs_hip import fr  # noqa: F401  registers the FR entry points
    from dgfs_hip.solvers.dgfsbi import scattering  # noqa: F401  registers the VSS entry points
    from dgfs_hip import surrogate  # noqa: F401  registers the surrogate entry points
    from dgfs_hip.integrators import graphs  # noqa: F401  registers the graph entry points
    from dgfs_hip.solvers.dgfsbi import system, initcond, inters  # noqa: F401
    from dgfs_hip import gmsh, native, nputil, plugins  # noqa: F401
    assert declared_symbols() <= set(_lib.exported_symbols())


def test_error_path_without_gpu_work():
    from dgfs_hip import _lib
    lib = _lib.load()
    # a null descriptor is rejected on the host before any device call
    st = lib.dgfs_vhs_create(None, None)
    assert st == 1
    assert b"null" in lib.dgfs_last_error()
    with pytest.raises(_lib.DGFSInvalidValue):
        _lib.check(st, "dgfs_vhs_create")
    assert lib.dgfs_vhs_destroy(None) == 5
    assert lib.dgfs_version().startswith(b"dgfs_hip")


def test_fr_entry_points_validate_on_host():
    """Every FR entry point converts its arguments and rejects null operands
    on the host, before any device work (no GPU needed)."""
    from dgfs_hip import _lib
    _lib.load()
    from dgfs_hip import fr
    N = None
    cases = [
        lambda: fr.mul(0, 4, 4, 4, 0, 8, 0, 8, 8, 1.0, 0.0),
        lambda: fr.tflux(2, 4, 4, 4, 0, 0, 128, 0, 128, 0, 128, 0, 8),
        lambda: fr.negdivconf(4, 4, 4, 0, 128, 0, 4),
        lambda: fr.intcflux(2, 4, 4, 0, 0, 0, 0, 0, 0, 0, 0, 8),
        lambda: fr.bccflux(0, 2, 4, 4, 0, 0, 0, 0, 0, N, N, N, (0, 0, 0), 0, 8),
        lambda: fr.bccupdate(0, 2, 4, 4, 0, 0, 0, 0, N, 0, (0, 0, 0), 1.0, 0, 8),
        lambda: fr.copy_reg(4, 4, 0, 128, 4, 0, 0, 128, 4, 0, 4),
        lambda: fr.moments(4, 4, 8, 0, 256, 0, 1.0, 0),
    ]
    for c in cases:
        with pytest.raises(_lib.DGFSInvalidValue):
            c()
    with pytest.raises(_lib.DGFSInvalidValue):
        _lib.call("frfs_axnpby", 10, 0, None, None, None)


def test_vss_entry_points_validate_on_host():
    """The multi-species operator rejects bad descriptors and handles on the
    host (no GPU needed)."""
    import ctypes as C
    import numpy as np
    from dgfs_hip import _lib
    from dgfs_hip.solvers.dgfsbi import scattering  # noqa: F401
    lib = _lib.load()
    assert lib.dgfs_vss_create(None, None) == 1
    a = np.zeros(64)
    dp = C.POINTER(C.c_double)
    ptr = a.ctypes.data_as(dp)
    for bad in (dict(nspcs=0), dict(nspcs=5), dict(Nv=10), dict(precision=16), dict(Mpre=0)):
        kw = dict(Nv=8, Nrho=2, M=6, Mpre=12, nspcs=2, precision=64, project=0, L=5.0,
                  sw=4 * np.pi / 6, cw=1.0, qz=ptr, qw=ptr, sz=ptr, szpre=ptr, masses=ptr,
                  gamma=ptr, eta=ptr, prefac=ptr)
        kw.update(bad)
        d = _lib.dgfs_vss_desc(**kw)
        h = C.c_void_p()
        st = lib.dgfs_vss_create(C.byref(d), C.byref(h))
        assert st in (1, 2), (bad, st)
        assert not h.value
    assert lib.dgfs_vss_destroy(None) == 5
    assert lib.dgfs_vss_eval_contig(None, None, None, 1, None) == 5
