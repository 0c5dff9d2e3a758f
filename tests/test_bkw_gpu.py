"""The homogeneous BKW relaxation (16^3 on [-6, 6]^3, M = 6, Nrho = 8, four
points) on the GPU, configured as a user would: Maxwell molecules
(omega = 1) at Kn = 1/Gamma(3/2), which the product's prefactor formula turns
into the BKW kernel 1/(4 pi); the initial state from DGFSBKWInitCondition
(t0 = 5.5); the reference's dgfs-tvd-rk2 / dgfs-euler steppers over the
homogeneous system.

Bars:
* FP64 trajectory vs the oracle's (golden file): <= 1e-10 relative L2;
* vs the analytic BKW solution at t = 6.5: < 2.5e-3 (the method's error at
  this resolution; the oracle's own trajectory is at 1.6e-3);
* conservation without projection: momentum <= 1e-14 (symmetry), mass and
  energy drift as the oracle's (3.0e-5, 3.9e-4) within 1e-9;
* FP32 transforms + FP64 projection: <= 1e-4 from the projected FP64 path,
  and mass, momentum and energy conserved to 1e-12 relative on both.
"""
import math
import os
import sys

import numpy as np
import pytest

from oracle import vhs as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "bkw_n16_steps.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(scheme, nsteps, precision="double", project=None, NvBatchSize=None):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import bkw_cfg
    from dgfs_hip.integrators.steppers import get_stepper
    from dgfs_hip.solvers.dgfs.homogeneous import DGFSHomogeneousSystem
    cfg = bkw_cfg(precision=precision, project=project, NvBatchSize=NvBatchSize)
    be = HIPBackend(cfg)
    nregs = {"dgfs-euler": 2, "dgfs-tvd-rk2": 3}[scheme]
    sysm = DGFSHomogeneousSystem(be, cfg, nregs)
    assert sysm.sm.prefactor == pytest.approx(1.0 / (4.0 * math.pi), rel=1e-15)
    st = get_stepper(scheme, be, sysm, 0.01, tstart=5.5)
    f0 = sysm.points(st.soln_bank)
    st.advance_to(5.5 + 0.01 * nsteps)
    assert st.nacptsteps == nsteps
    return sysm, st, f0, sysm.points(st.soln_bank)


def test_bkw_initial_state_is_the_analytic_one(device):
    sysm, st, f0, _ = run("dgfs-euler", 0)
    vm = O.VelocityMesh(16, 6.0, 8, 6)
    assert np.allclose(f0, O.bkw(vm, 5.5)[None], rtol=1e-14, atol=1e-300)
    mom = sysm.moments(st.soln_bank)
    # BKW: rho = 1, U = 0, T = 2 in DGFS units (<c^2> = 3)
    assert np.allclose(mom[:, 0], 1.0, rtol=1e-6)
    assert np.abs(mom[:, 1:4]).max() < 1e-14
    assert np.allclose(mom[:, 4], 2.0, rtol=1e-3)


def test_bkw_q_from_kn_config_is_the_time_derivative(device):
    """One collide of the configured system on the BKW state against the
    oracle's Q (1e-10) and the closed-form dK/dt derivative (the discretisation
    error at 16^3 / M = 6, 1.8e-2)."""
    sysm, st, f0, _ = run("dgfs-euler", 0)
    sysm.collide(5.5, 0, 1)
    Q = sysm.points(1)
    vm = O.VelocityMesh(16, 6.0, 8, 6)
    Qo = O.VHSOracle(vm, 0.0, 1.0 / (4.0 * math.pi)).q(f0[0])
    e = rel(Q, np.broadcast_to(Qo, Q.shape))
    say("Q vs oracle", e)
    assert e < 1e-10
    ed = rel(Q[0], O.bkw_dfdt(vm, 5.5))
    say("Q vs dfdt", ed)
    assert ed < 2.5e-2


def test_bkw_relaxation_rk2_100_steps(device):
    g = np.load(GOLD)
    sysm, st, f0, f = run("dgfs-tvd-rk2", 100)
    vm = O.VelocityMesh(16, 6.0, 8, 6)
    e_or = max(rel(f[i], g["f_rk2"]) for i in range(4))
    fa = O.bkw(vm, 6.5)
    e_an = max(rel(f[i], fa) for i in range(4))
    say("rk2 100 steps: vs oracle", e_or, "vs analytic", e_an)
    assert e_or < 1e-10
    assert e_an < 2.5e-3
    assert np.abs(f - f[:1]).max() == 0.0 or rel(f, np.broadcast_to(f[:1], f.shape)) < 1e-14
    i0 = O.collision_invariants(vm, f0)
    i1 = O.collision_invariants(vm, f)
    io = O.collision_invariants(vm, g["f_rk2"])
    assert np.abs(i1[:, 1:4]).max() < 1e-14
    assert np.allclose(i1[:, 0] - i0[:, 0], io[0] - i0[0, 0], rtol=1e-6, atol=1e-12)
    assert np.allclose(i1[:, 4] - i0[:, 4], io[4] - i0[0, 4], rtol=1e-6, atol=1e-12)
    mom = sysm.moments(st.soln_bank)
    say("moments t=6.5", mom[0, :5])


def test_bkw_euler_10_steps_batched_velocities(device):
    """The Euler stepper with the velocities in 4 batches (NvBatchSize = 1024)."""
    g = np.load(GOLD)
    _, _, _, f = run("dgfs-euler", 10, NvBatchSize=1024)
    e = max(rel(f[i], g["f_euler"]) for i in range(4))
    say("euler 10 steps vs oracle", e)
    assert e < 1e-10


def test_bkw_relaxation_fp32_projected(device):
    _, _, f0, f64p = run("dgfs-tvd-rk2", 100, project=True)
    _, _, _, f32 = run("dgfs-tvd-rk2", 100, precision="single", project=True)
    vm = O.VelocityMesh(16, 6.0, 8, 6)
    e = rel(f32, f64p)
    e_an = rel(f32[0], O.bkw(vm, 6.5))
    say("fp32+projection vs fp64+projection", e, "vs analytic", e_an)
    assert e < 1e-4
    assert e_an < 2.5e-3
    i0 = O.collision_invariants(vm, f0)
    scale = np.array([i0[0, 0], 1.0, 1.0, 1.0, i0[0, 4]])
    for f in (f64p, f32):
        drift = np.abs(O.collision_invariants(vm, f) - i0) / scale
        say("invariant drift", drift.max(axis=0))
        assert drift.max() < 1e-12
