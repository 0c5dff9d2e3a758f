"""Checkpoint and restart on the GPU: a run that writes solution files and a
second run started from one of them end in bitwise the same registers as an
uninterrupted run (single and two-species systems), and a solution of a
lower order restarts a higher-order system through the basis interpolation."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _single(initsoln=None, order=2):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=order, uwall=0.3)
    be = HIPBackend(cfg)
    mesh = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    return be, DGFSSystem(be, mesh, cfg, 3, initsoln=initsoln)


def _bi(initsoln=None):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import bi_couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfsbi.system import DGFSBiSystem
    cfg = bi_couette_cfg(Nv=16, L=4.0, Nrho=4, M=12, Mpre=24, masses=(1.0, 2.1), nden=(0.6, 0.4),
                         omega=0.5, alpha=1.2, prefactor=0.05, order=1, uwall=0.3)
    be = HIPBackend(cfg)
    mesh = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    return be, DGFSBiSystem(be, mesh, cfg, 6, initsoln=initsoln)


@pytest.mark.parametrize("kind,stepper", [("dgfs", "dgfs-tvd-rk2"), ("dgfsbi", "dgfsbi-tvd-rk2")])
def test_restart_is_bitwise(kind, stepper, tmp_path, device):
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.native import NativeReader
    from dgfs_hip.plugins import DGFSDistWriterStdPlugin
    build = _single if kind == "dgfs" else _bi
    dt = 0.01
    be, sysm = build()
    st = get_stepper(stepper, be, sysm, dt)
    wr = DGFSDistWriterStdPlugin(st, 2 * dt, str(tmp_path), "run-{n:02d}")
    st.plugins.append(wr)
    say("restart", kind, "uninterrupted 4 steps")
    st.advance_to(4 * dt)
    full = sysm.soln_arrays(st.soln_bank)
    assert [p.rsplit("/", 1)[1] for p in wr.paths] == ["run-00.frfss", "run-01.frfss", "run-02.frfss"]
    r = NativeReader(wr.paths[1])
    be2, sys2 = build(initsoln=r)
    st2 = get_stepper(stepper, be2, sys2, dt)
    st2.restart(r)
    assert st2.nacptsteps == 2 and st2.tcurr == 2 * dt
    wr2 = DGFSDistWriterStdPlugin(st2, 2 * dt, str(tmp_path), "run-{n:02d}", isrestart=True)
    st2.plugins.append(wr2)
    say("restart", kind, "from step 2")
    st2.advance_to(4 * dt)
    again = sys2.soln_arrays(st2.soln_bank)
    for et in full:
        assert np.array_equal(full[et], again[et]), np.abs(full[et] - again[et]).max()
    assert st2.nacptsteps == 4
    assert [p.rsplit("/", 1)[1] for p in wr2.paths] == ["run-03.frfss"]
    assert np.array_equal(NativeReader(wr2.paths[0])["soln_quad_p0"], NativeReader(wr.paths[2])["soln_quad_p0"])


def test_restart_raises_order(tmp_path, device):
    """p = 1 solution into a p = 2 system: the registers hold the p = 1
    field evaluated at the p = 2 points."""
    from dgfs_hip.native import NativeReader, write_native
    from dgfs_hip.shapes import get_shape
    be, s1 = _single(order=1)
    rng = np.random.default_rng(3)
    f1 = s1.eles_scal_upts_full[0].get() * (1 + 0.1 * rng.random((4, 1, 6)))
    s1.eles_scal_upts_full[0].set(f1)
    from dgfs_hip.inifile import Inifile
    stats = Inifile()
    stats.set("data", "fields", s1.dist_fields())
    stats.set("solver-time-integrator", "tcurr", "0.0")
    p = write_native(str(tmp_path / "p1.frfss"), {"soln_quad_p0": s1.soln_arrays(0)["quad"]},
                     {"config": s1.cfg.tostr(), "stats": stats.tostr(), "mesh_uuid": s1.mesh.uuid})
    say("restart p=1 -> p=2")
    be2, s2 = _single(initsoln=NativeReader(p), order=2)
    got = s2.eles_scal_upts_full[0].get()
    B = get_shape("quad", 1, "gauss-legendre").ubasis.at(get_shape("quad", 2, "gauss-legendre").upts)
    want = np.einsum("ij,jve->ive", B, f1)
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()


def test_restart_rejects_other_mesh(tmp_path, device):
    from dgfs_hip.native import NativeReader, write_native
    be, s1 = _single()
    from dgfs_hip.inifile import Inifile
    stats = Inifile()
    stats.set("data", "fields", s1.dist_fields())
    p = write_native(str(tmp_path / "m.frfss"), {"soln_quad_p0": s1.soln_arrays(0)["quad"]},
                     {"config": s1.cfg.tostr(), "stats": stats.tostr(), "mesh_uuid": "another"})
    with pytest.raises(RuntimeError, match="mesh"):
        _single(initsoln=NativeReader(p))
