"""HIP-graph replay of DGFS time steps: every replayed step equals the eager
step (same kernels, same inputs), with eager work and synchronisations
between the replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _system(own_stream, NvBS):
    from dgfs_hip.backend import HIPBackend
    from dgfs_hip.cases import couette_cfg
    from dgfs_hip.mesh import rect_quads
    from dgfs_hip.solvers.dgfs.system import DGFSSystem
    cfg = couette_cfg(Nv=16, L=5.0, Nrho=4, M=6, Kn=1.0, omega=0.5, order=2, uwall=0.3, NvBatchSize=NvBS)
    be = HIPBackend(cfg, own_stream=own_stream)
    mesh = rect_quads(3, 2, 0.0, 1.5, 0.0, 1.0, periodic_x=True)
    return be, DGFSSystem(be, mesh, cfg, 3)


@pytest.mark.parametrize("stepper,nbs", [("dgfs-tvd-rk2", 1024), ("dgfs-euler", None)])
def test_graph_replay_matches_eager(stepper, nbs, device):
    import torch
    from dgfs_hip.integrators.graphs import GraphStepper
    from dgfs_hip.integrators.steppers import get_stepper
    be0, s0 = _system(False, nbs)
    be1, s1 = _system(True, nbs)
    rng = np.random.default_rng(5)
    f0 = s0.eles_scal_upts_full[0].get()
    f0 = f0 * (1 + 0.1 * rng.random((f0.shape[0], 1, f0.shape[2])))
    s0.eles_scal_upts_full[0].set(f0)
    s1.eles_scal_upts_full[0].set(f0)
    dt = 0.01
    eager = get_stepper(stepper, be0, s0, dt)
    gs = GraphStepper(get_stepper(stepper, be1, s1, dt))
    scratch = torch.zeros(1 << 20, device=device)
    for k in range(5):
        say("step", k, stepper)
        eager._idxcurr = eager.step(eager.tcurr, dt)
        eager.tcurr += dt
        gs.stepper._idxcurr = gs.step(gs.stepper.tcurr, dt)
        gs.stepper.tcurr += dt
        # eager work of the program between replays
        scratch.add_(1.0)
        torch.cuda.synchronize()
        a = s0.eles_scal_upts_full[eager.soln_bank].get()
        b = s1.eles_scal_upts_full[gs.stepper.soln_bank].get()
        err = float(np.abs(a - b).max() / np.abs(a).max())
        say("  max rel diff", err, "graph nodes", gs.nodes)
        assert err <= 1e-15
    assert gs.nodes > 10
    assert float(scratch[0]) == 5.0
    gs.close()


@pytest.mark.parametrize("stepper", ["dgfs-tvd-rk2", "dgfs-euler"])
def test_graph_advance_runs_plugins_and_restarts(stepper, device):
    """GraphStepper.advance_to calls every plugin after each accepted step,
    counts the stepper's right-hand-side evaluations, and a stepper that
    resumes at a non-zero step count (a restart) still warms up eagerly before
    the capture; its states equal the eager stepper's."""
    from dgfs_hip.integrators.graphs import GraphStepper
    from dgfs_hip.integrators.steppers import get_stepper
    be0, s0 = _system(False, None)
    be1, s1 = _system(True, None)
    dt = 0.01
    eager = get_stepper(stepper, be0, s0, dt)
    inner = get_stepper(stepper, be1, s1, dt)
    inner.tcurr, inner.nacptsteps = 0.3, 30          # as restart() leaves them
    eager.tcurr, eager.nacptsteps = 0.3, 30
    seen = []
    inner.plugins.append(lambda s: seen.append((s.nacptsteps, s.nfevals)))
    gs = GraphStepper(inner)
    say("graph advance", stepper)
    gs.advance_to(0.3 + 4 * dt)
    eager.advance_to(0.3 + 4 * dt)
    per = 2 if stepper == "dgfs-tvd-rk2" else 1
    assert seen == [(31 + k, per * (k + 1)) for k in range(4)]
    assert inner.nfevals == eager.nfevals == 4 * per
    a = s0.eles_scal_upts_full[eager.soln_bank].get()
    b = s1.eles_scal_upts_full[inner.soln_bank].get()
    assert float(np.abs(a - b).max() / np.abs(a).max()) <= 1e-15
    assert gs.nodes > 10
    gs.close()
