"""The allocation-free Q entry point (dgfs_vhs_eval_contig_noalloc) against
the allocating one (dgfs_vhs_eval_contig), through the C-ABI, on two calls in
a row with the second batch smaller.  The two differ only in how points are
chunked and split into radial groups, i.e. in the summation order over the
radial nodes, so the bars are roundoff-level: FP64 <= 1e-12 relative L2
(inside the 1e-10 FP64 parity bar) and FP32 <= 1e-4 (two FP32 evaluations,
each about 3e-5 from FP64 after the gain/loss cancellation; measured 1.6e-5).
The FP64 case is also held to the FP64 oracle at 1e-10, and the FP32 cases to
the FP64 operator with the same projection at the FP32 bar, 1e-3.  Device
free memory is the same before and after the allocation-free calls."""
import sys

import numpy as np
import pytest

from oracle import vhs as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def states(vm, npts, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(npts):
        u1, u2 = rng.uniform(-0.6, 0.6, 3), rng.uniform(-0.6, 0.6, 3)
        t1, t2 = rng.uniform(0.5, 2.0, 2)
        f = 0.5 * O.maxwellian(vm, 1, u1, t1) + 0.5 * O.maxwellian(vm, 1, u2, t2)
        out.append(f / (f.sum() * vm.cw))
    return np.array(out)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [
    # Nv, L, Nrho, M, precision, project, (first, second) batch sizes, bar
    (16, 6.0, 8, 6, 64, False, (40, 7), 1e-12),
    (32, 6.0, 16, 12, 32, True, (300, 45), 1e-4),
    (24, 6.0, 16, 12, 32, True, (200, 30), 1e-4),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"nv{c[0]}_fp{c[4]}")
def test_noalloc_matches_allocating_entry(case, device):
    import torch
    from dgfs_hip.solvers.dgfs.scattering import VHSOperator
    from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
    N, L, Nrho, M, prec, project, sizes, bar = case
    vm = DGFSVelocityMesh.from_params(None, N, L, Nrho, M)
    ov = O.VelocityMesh(N, L, Nrho, M)
    op = VHSOperator.from_mesh(vm, 0.38, 1.0, precision=prec, project=project)
    stream = torch.cuda.current_stream().cuda_stream
    fs = [torch.as_tensor(states(ov, n, seed=40 + i), device=device) for i, n in enumerate(sizes)]
    Qs = [torch.full_like(f, float("nan")) for f in fs]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for f, Q in zip(fs, Qs):
        say(f"noalloc Nv={N} precision={prec} npts={f.shape[0]}")
        op.eval_contig_noalloc(f.data_ptr(), Q.data_ptr(), f.shape[0], stream)
        torch.cuda.synchronize()
        free = torch.cuda.mem_get_info()[0]
        say("  device free bytes before / after:", free0, free)
        assert free == free0
    for f, Q in zip(fs, Qs):
        ref = torch.full_like(f, float("nan"))
        say(f"allocating entry Nv={N} precision={prec} npts={f.shape[0]}")
        op.eval_contig(f.data_ptr(), ref.data_ptr(), f.shape[0], stream)
        torch.cuda.synchronize()
        q, r = Q.cpu().numpy(), ref.cpu().numpy()
        assert np.isfinite(q).all()
        e = rel(q, r)
        say("  rel L2 (noalloc vs allocating):", e)
        assert e < bar
    if prec == 64:
        f = fs[1].cpu().numpy()
        e = rel(Qs[1].cpu().numpy(), O.VHSOracle(ov, 0.38, 1.0).q(f))
        say("  FP64 noalloc vs oracle:", e)
        assert e < 1e-10
    else:
        op64 = VHSOperator.from_mesh(vm, 0.38, 1.0, precision=64, project=project)
        ref = torch.full_like(fs[1], float("nan"))
        op64.eval_contig(fs[1].data_ptr(), ref.data_ptr(), fs[1].shape[0], stream)
        torch.cuda.synchronize()
        e = rel(Qs[1].cpu().numpy(), ref.cpu().numpy())
        say("  FP32 noalloc vs FP64 (same projection):", e)
        assert e < 1e-3


def test_noalloc_argument_errors(device):
    from dgfs_hip import _lib
    from dgfs_hip.solvers.dgfs.scattering import VHSOperator
    from dgfs_hip.solvers.dgfs.velocitymesh import DGFSVelocityMesh
    vm = DGFSVelocityMesh.from_params(None, 8, 5.0, 4, 6)
    op = VHSOperator.from_mesh(vm, 0.0, 1.0)
    with pytest.raises(_lib.DGFSError):
        op.eval_contig_noalloc(0, 0, 4)
    with pytest.raises(_lib.DGFSError):
        op.eval_contig_noalloc(1 << 20, 1 << 20, 4)
    op.eval_contig_noalloc(1 << 20, 2 << 20, 0)  # an empty batch touches nothing
