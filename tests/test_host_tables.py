"""Host-side tables of the product (quadrature, designs, velocity mesh) agree
with the oracle's restatement.  CPU only."""
import numpy as np
import pytest

from oracle import vhs as O


@pytest.mark.parametrize("N,L,Nrho,M", [(8, 5.0, 4, 6), (24, 7.0, 12, 12), (32, 6.0, 16, 12), (16, 6.0, 8, 3)])
def test_velocity_mesh_matches_oracle(N, L, Nrho, M):
    from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
    vm = DGFSVelocityMesh.from_params(None, N, L, Nrho, M)
    ov = O.VelocityMesh(N, L, Nrho, M)
    assert vm.L() == pytest.approx(L, rel=1e-15)
    assert np.allclose(vm.cv(), ov.cv, rtol=0, atol=1e-15)
    assert vm.cw() == pytest.approx(ov.cw, rel=1e-15)
    assert np.allclose(vm.qz(), ov.qz, rtol=1e-14) and np.allclose(vm.qw(), ov.qw, rtol=1e-13)
    assert np.allclose(vm.sz(), ov.sz, atol=1e-15)
    assert vm.sw() == pytest.approx(ov.sw, rel=1e-15)


def test_lobatto_rule_matches_oracle():
    from dgfs_hip.quadratures import zwglj
    for n in (2, 3, 5, 9):
        z, w = zwglj(n)
        z2, w2 = O.gauss_lobatto_jacobi(n)
        assert np.allclose(z, z2, atol=1e-14) and np.allclose(w, w2, atol=1e-14)


def test_unknown_design_raises():
    from dgfs_hip.sphericaldesign import get_sphquadrule
    with pytest.raises(ValueError):
        get_sphquadrule("symmetric", rule="womersley", npts=14)


def test_inline_design_rule():
    from dgfs_hip.sphericaldesign import get_sphquadrule
    r = get_sphquadrule("symmetric", rule="0 0 1\n0 0 -1\n", npts=2)
    assert r.pts.shape == (2, 3)


def test_vm_batch_checks():
    from dgfs_hip.inifile import Inifile
    from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
    vm = DGFSVelocityMesh.from_params(None, 8, 5.0, 4, 6)
    cfg = vm.cfg
    cfg.set("constants", "NvBatchSize", 7)
    with pytest.raises(ValueError):
        DGFSVelocityMesh(None, cfg)
    assert isinstance(cfg, Inifile)
