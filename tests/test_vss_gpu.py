"""GPU parity of the multi-species VSS collision operator against the FP64
oracle (oracle/vss.py) and the committed golden vectors.  Tolerances: FP64
path <= 1e-10 relative L2 per species; FP32 path <= 1e-3; after the FP64
projection the species masses, total momentum and total energy of Q vanish
to <= 1e-12 of int |Q|."""
import os
import sys

import numpy as np
import pytest

from oracle import vss as V

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
PAIRS = [(0, 0), (0, 1), (1, 0), (1, 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_op(N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac, precision=64, project=False):
    from dgfs_hip.solvers.dgfsbi.scattering import VSSOperator
    from dgfs_hip.solvers.dgfsbi.velocitymesh import DGFSBiVelocityMesh
    from dgfs_hip.sphericaldesign import get_sphquadrule
    vm = DGFSBiVelocityMesh.from_params(None, N, L, Nrho, M, masses)
    S = len(masses)
    gamma = np.array([2.0 * (1.0 - w) for w in omegas]).reshape(S, S)
    eta = np.full((S, S), alpha - 1.0)
    szpre = get_sphquadrule("symmetric", npts=Mpre).pts
    op = VSSOperator(N, Nrho, M, vm.L(), vm.sw(), vm.cw(), vm.qz(), vm.qw(), vm.sz(), szpre,
                     vm.masses(), gamma, eta, np.reshape(prefac, (S, S)), precision=precision,
                     project=project)
    return op, vm


def run_contig(op, fs, device):
    import torch
    ft = [torch.as_tensor(np.ascontiguousarray(f), dtype=torch.float64, device=device) for f in fs]
    Qt = [torch.full_like(f, float("nan")) for f in ft]
    op.eval_contig([f.data_ptr() for f in ft], [q.data_ptr() for q in Qt], ft[0].shape[0],
                   torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return [q.cpu().numpy() for q in Qt]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_q(N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac, fs):
    vm = V.BiVelocityMesh(N, L, Nrho, M, masses)
    g = {k: 2.0 * (1.0 - w) for k, w in zip(PAIRS, omegas)}
    eta = {k: alpha - 1.0 for k in PAIRS}
    orc = V.VSSOracle(vm, g, eta, dict(zip(PAIRS, prefac)), V.full_sphere_design(Mpre))
    return orc.q(fs), vm


def bimodal_states(vm, npts, seed):
    rng = np.random.default_rng(seed)
    out = [[], []]
    for _ in range(npts):
        for p in range(2):
            u1, u2 = rng.uniform(-0.6, 0.6, 3), rng.uniform(-0.6, 0.6, 3)
            t1, t2 = rng.uniform(0.6, 1.8, 2)
            out[p].append(0.5 * V.species_maxwellian(vm, p, 1.0, u1, t1)
                          + 0.5 * V.species_maxwellian(vm, p, 0.6, u2, t2))
    return [np.array(o) for o in out]


@pytest.mark.parametrize("fn", sorted(f for f in os.listdir(GOLD) if f.startswith("vss_") and f.endswith(".npz")))
def test_golden_vectors_fp64(fn, device):
    d = np.load(os.path.join(GOLD, fn))
    N, Nrho, M, Mpre = (int(x) for x in d["dims"])
    say("golden", fn, N, Nrho, M, Mpre)
    op, _ = make_op(N, float(d["L"]), Nrho, M, Mpre, tuple(d["masses"]), float(d["alpha"]),
                    tuple(d["omegas"]), tuple(d["prefac"]))
    Q0, Q1 = run_contig(op, [d["f0"], d["f1"]], device)
    e0, e1 = rel(Q0, d["Q0"]), rel(Q1, d["Q1"])
    say("  rel L2", e0, e1)
    assert e0 < 1e-10 and e1 < 1e-10


CFG16 = (16, 5.0, 8, 12, 24, (1.0, 2.1), 1.0, (0.81, 0.805, 0.805, 0.80), (1.0, 0.9, 0.9, 0.8))


def test_oracle_parity_fp64_fp32_16(device):
    N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac = CFG16
    vm = V.BiVelocityMesh(N, L, Nrho, M, masses)
    fs = bimodal_states(vm, 3, seed=5)
    Qref, _ = oracle_q(*CFG16, fs)
    say("vss oracle parity 16^3")
    op64, _ = make_op(*CFG16, precision=64)
    Q64 = run_contig(op64, fs, device)
    e64 = [rel(Q64[p], Qref[p]) for p in range(2)]
    say("  fp64", e64)
    assert max(e64) < 1e-10
    op32, _ = make_op(*CFG16, precision=32)
    Q32 = run_contig(op32, fs, device)
    e32 = [rel(Q32[p], Qref[p]) for p in range(2)]
    say("  fp32", e32)
    assert max(e32) < 1e-3


# the cylinder workload's grid and rule at a reduced radial rule: 24^3 on
# [-6, 6]^3, M = 12, Nrho = 6, Mpre = 24, VHS (alpha = 1, the isotropic gain)
# and VSS (alpha = 1.3, the general gain)
CFG24 = (24, 6.0, 6, 12, 24, (1.0, 2.1), 1.0, (0.81, 0.805, 0.805, 0.80), (1.0, 0.9, 0.9, 0.8))


@pytest.mark.parametrize("alpha", [1.0, 1.3])
def test_oracle_parity_fp64_fp32_24(alpha, device):
    """24^3 against the FP64 CPU oracle on 2 bimodal mixture points: the FP64
    operator (the slab kernel for alpha = 1, the multi-pass one for the
    general gain) to 1e-10 and the FP32 one (register-resident, the default
    at 24^3) to 1e-3."""
    cfg = CFG24[:6] + (alpha,) + CFG24[7:]
    N, L, Nrho, M, Mpre, masses = cfg[:6]
    vm = V.BiVelocityMesh(N, L, Nrho, M, masses)
    fs = bimodal_states(vm, 2, seed=24)
    Qref, _ = oracle_q(*cfg, fs)
    say("vss oracle parity 24^3, alpha", alpha)
    op64, _ = make_op(*cfg, precision=64)
    Q64 = run_contig(op64, fs, device)
    e64 = [rel(Q64[p], Qref[p]) for p in range(2)]
    say("  fp64 algorithm", op64.get_algorithm(), e64)
    assert op64.get_algorithm() == (7 if alpha == 1.0 else 1)
    assert max(e64) < 1e-10
    op32, _ = make_op(*cfg, precision=32)
    Q32 = run_contig(op32, fs, device)
    say("  fp32 algorithm", op32.get_algorithm())
    e32 = [rel(Q32[p], Qref[p]) for p in range(2)]
    say("  fp32", e32)
    assert op32.get_algorithm() >= 4
    assert max(e32) < 1e-3


def test_group_split_and_aosoa_path(device):
    import torch
    N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac = 8, 5.0, 4, 12, 24, (1.0, 2.1), 1.2, \
        (0.81, 0.805, 0.805, 0.80), (1.0, 1.0, 1.0, 1.0)
    op, vm = make_op(N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac)
    ov = V.BiVelocityMesh(N, L, Nrho, M, masses)
    nupts, neles = 4, 37
    fs = bimodal_states(ov, nupts * neles, seed=11)
    op.set_groups(1)
    Qa = run_contig(op, fs, device)
    op.set_groups(Nrho)
    Qb = run_contig(op, fs, device)
    say("group split", [rel(Qa[p], Qb[p]) for p in range(2)])
    assert all(rel(Qa[p], Qb[p]) < 1e-13 for p in range(2))
    op.set_groups(0)
    # AoSoA registers: point pt = (elem - e0) * nupts + upt
    n = N ** 3
    ldim = ((neles + 31) // 32) * n * 32
    regs, outs = [], []
    for p in range(2):
        R = torch.zeros((nupts, ldim), dtype=torch.float64, device=device)
        for e in range(neles):
            for u in range(nupts):
                col = ((e // 32) * n + np.arange(n)) * 32 + e % 32
                R[u, torch.as_tensor(col, device=device)] = torch.as_tensor(fs[p][e * nupts + u], device=device)
        regs.append(R)
        outs.append(torch.full_like(R, 7.0))
    e0, ne = 3, 30
    op.eval_aosoa([r.data_ptr() for r in regs], [o.data_ptr() for o in outs], nupts, ldim, n, neles, e0, ne,
                  torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    Qc = run_contig(op, fs, device)
    for p in range(2):
        O = outs[p].cpu().numpy()
        for e in range(neles):
            for u in range(nupts):
                col = ((e // 32) * n + np.arange(n)) * 32 + e % 32
                if e0 <= e < e0 + ne:
                    assert rel(O[u, col], Qc[p][e * nupts + u]) < 1e-14
                else:
                    assert np.all(O[u, col] == 7.0)


def test_projection_conserves_mixture_invariants(device):
    N, L, Nrho, M, Mpre, masses, alpha, omegas, prefac = CFG16
    vm = V.BiVelocityMesh(N, L, Nrho, M, masses)
    fs = bimodal_states(vm, 4, seed=9)
    op64p, _ = make_op(*CFG16, precision=64, project=True)
    op32p, _ = make_op(*CFG16, precision=32, project=True)
    Q64 = run_contig(op64p, fs, device)
    Q32 = run_contig(op32p, fs, device)
    cv, cw, m = vm.cv, vm.cw, vm.masses
    for Q in (Q64, Q32):
        scale = sum(np.abs(Q[p]).sum(1) for p in range(2)) * cw
        for p in range(2):
            assert np.max(np.abs(Q[p].sum(1) * cw) / scale) < 1e-12
        mom = sum(m[p] * (Q[p] @ cv.T) for p in range(2)) * cw
        en = sum(m[p] * (Q[p] @ (cv ** 2).sum(0)) for p in range(2)) * cw
        assert np.max(np.abs(mom) / scale[:, None]) < 1e-12
        assert np.max(np.abs(en) / scale) < 1e-12
    e = [rel(Q32[p], Q64[p]) for p in range(2)]
    say("fp32 vs fp64 (projected)", e)
    assert max(e) < 1e-3


def test_bad_arguments_raise(device):
    import torch
    from dgfs_hip import _lib
    op, _ = make_op(4, 5.0, 2, 6, 12, (1.0, 2.1), 1.0, (0.8, 0.8, 0.8, 0.8), (1, 1, 1, 1))
    f = torch.zeros((2, 64), dtype=torch.float64, device=device)
    with pytest.raises(_lib.DGFSInvalidValue):
        op.eval_contig([f.data_ptr(), f.data_ptr()], [f.data_ptr(), f.data_ptr()], 2)
    with pytest.raises(ValueError):
        op.eval_contig([f.data_ptr()], [f.data_ptr()], 2)


def _mixture_states(ov, npts, seed):
    rng = np.random.default_rng(seed)
    out = []
    for p in range(len(ov.masses)):
        rows = []
        for _ in range(npts):
            u1, u2 = rng.uniform(-0.5, 0.5, 3), rng.uniform(-0.5, 0.5, 3)
            t1, t2 = rng.uniform(0.7, 1.6, 2)
            rows.append(0.6 * V.species_maxwellian(ov, p, 1.0, u1, t1) + 0.4 * V.species_maxwellian(ov, p, 1.0, u2, t2))
        out.append(np.array(rows))
    return out


@pytest.mark.parametrize("masses,precision,tol", [((1.0, 2.1), 64, 1e-13), ((1.0, 2.1, 0.7), 64, 1e-13),
                                                  ((1.0, 2.1), 32, 1e-5)])
def test_antipodal_sharing_matches_pair_loop(masses, precision, tol, device):
    """One product per (x, y) for both (p, q) and (q, p) at -y, and half the
    directions with summed antipodal rows for (p, p), against the
    reference's per-pair loop over every direction (same handle)."""
    S = len(masses)
    N, L, Nrho, M, Mpre = 8, 5.0, 4, 12, 24
    omegas = tuple(0.8 + 0.005 * (p + q) for p in range(S) for q in range(S))
    prefac = tuple(1.0 + 0.1 * k for k in range(S * S))
    op, _ = make_op(N, L, Nrho, M, Mpre, masses, 1.2, omegas, prefac, precision=precision)
    assert op.get_symmetry()
    fs = _mixture_states(V.BiVelocityMesh(N, L, Nrho, M, masses), 6, seed=21)
    Qs = run_contig(op, fs, device)
    op.set_symmetry(False)
    assert not op.get_symmetry()
    Qp = run_contig(op, fs, device)
    e = [rel(Qs[p], Qp[p]) for p in range(S)]
    say("sharing vs pair loop", S, precision, e)
    assert max(e) < tol


def test_sharing_needs_antipodes(device):
    from dgfs_hip import _lib
    from dgfs_hip.solvers.dgfsbi.scattering import VSSOperator
    from dgfs_hip.solvers.dgfsbi.velocitymesh import DGFSBiVelocityMesh
    from dgfs_hip.sphericaldesign import get_sphquadrule
    N, L, Nrho, M = 4, 5.0, 2, 6
    vm = DGFSBiVelocityMesh.from_params(None, N, L, Nrho, M, (1.0, 2.1))
    sz = np.array(vm.sz()).reshape(M, 3).copy()
    sz[0] = np.array([0.6, 0.8, 0.0])       # no longer the negation of its partner
    op = VSSOperator(N, Nrho, M, vm.L(), vm.sw(), vm.cw(), vm.qz(), vm.qw(), sz,
                     get_sphquadrule("symmetric", npts=12).pts, vm.masses(), np.full((2, 2), 0.4),
                     np.zeros((2, 2)), np.ones((2, 2)), precision=64)
    assert not op.get_symmetry()
    with pytest.raises(_lib.DGFSUnsupported):
        op.set_symmetry(True)


@pytest.mark.parametrize("N,L,masses,sym,alpha", [(16, 5.0, (1.0, 2.1), True, 1.0), (24, 6.0, (1.0, 2.1), True, 1.0),
                                                  (24, 6.0, (1.0, 2.1), False, 1.0),
                                                  (16, 5.0, (1.0, 2.1, 0.7), True, 1.0),
                                                  (24, 6.0, (1.0, 2.1), True, 1.3),
                                                  (16, 5.0, (1.0, 2.1), False, 1.3)])
def test_register_resident_matches_multipass(N, L, masses, sym, alpha, device):
    """The register-resident FP32 gain (algorithm 4) and the multi-pass FP32
    gain on the same handle and states, both against the FP64 operator: the
    same sums with other FP32 roundings, so the register-resident error is
    within 1e-4 and within twice the multi-pass error, and the two FP32
    results agree to 5e-5.  alpha = 1 (eta = 0) takes the isotropic kernel,
    which sums the direction products before one forward transform per
    radial node; alpha = 1.3 the general one."""
    S = len(masses)
    Nrho, M, Mpre = 6, 12, 24
    omegas = tuple(0.81 - 0.005 * (p + q) for p in range(S) for q in range(S))
    op, _ = make_op(N, L, Nrho, M, Mpre, masses, alpha, omegas, (1.0,) * (S * S), precision=32)
    op64, _ = make_op(N, L, Nrho, M, Mpre, masses, alpha, omegas, (1.0,) * (S * S), precision=64)
    op.set_symmetry(sym)
    fs = _mixture_states(V.BiVelocityMesh(N, L, Nrho, M, masses), 5, seed=31)
    Q64 = run_contig(op64, fs, device)
    op.set_algorithm(4)
    Qr = run_contig(op, fs, device)
    assert op.get_algorithm() == (5 if alpha == 1.0 else 4)
    op.set_algorithm(1)
    Qm = run_contig(op, fs, device)
    assert op.get_algorithm() == 1
    er = [rel(Qr[p], Q64[p]) for p in range(S)]
    em = [rel(Qm[p], Q64[p]) for p in range(S)]
    e = [rel(Qr[p], Qm[p]) for p in range(S)]
    say("register-resident / multi-pass vs FP64, and between", N, S, sym, er, em, e)
    for p in range(S):
        assert er[p] < 1e-4 and er[p] < 2 * em[p] + 1e-6
    assert max(e) < 5e-5
    if alpha == 1.0:
        # the packed kernel: one complex inverse transform for t3 and t4 plus
        # the Nyquist-plane corrections
        op.set_algorithm(6)
        Qk = run_contig(op, fs, device)
        assert op.get_algorithm() == 6
        ek = [rel(Qk[p], Q64[p]) for p in range(S)]
        say("  packed vs FP64", ek)
        for p in range(S):
            assert ek[p] < 1e-4 and ek[p] < 2 * em[p] + 1e-6
    else:
        from dgfs_hip import _lib
        with pytest.raises(_lib.DGFSUnsupported):
            op.set_algorithm(6)


@pytest.mark.parametrize("N,L,masses,sym", [(16, 5.0, (1.0, 2.1), True), (24, 6.0, (1.0, 2.1), True),
                                            (24, 6.0, (1.0, 2.1), False), (16, 5.0, (1.0, 2.1, 0.7), True)])
def test_slab_kernel_matches_multipass(N, L, masses, sym, device):
    """The slab gain for isotropic scattering (algorithm 7: x-slabs of t3 and
    t4 in LDS, S_x in registers, x-forward sum in a second kernel), the FP64
    default at Nv 16 / 24, against the multi-pass FP64 gain on the same
    states (roundoff) and, in FP32, against the FP64 operator."""
    S = len(masses)
    Nrho, M, Mpre = 6, 12, 24
    omegas = tuple(0.81 - 0.005 * (p + q) for p in range(S) for q in range(S))
    op64, _ = make_op(N, L, Nrho, M, Mpre, masses, 1.0, omegas, (1.0,) * (S * S), precision=64)
    op64.set_symmetry(sym)
    fs = _mixture_states(V.BiVelocityMesh(N, L, Nrho, M, masses), 4, seed=37)
    op64.set_algorithm(1)
    Qm = run_contig(op64, fs, device)
    for groups in (0, 1, 4):
        say("vss fp64 slab kernel", N, S, sym, "groups", groups)
        op64.set_algorithm(7)
        op64.set_groups(groups)
        Qs = run_contig(op64, fs, device)
        assert op64.get_algorithm() == 7
        d = [rel(Qs[p], Qm[p]) for p in range(S)]
        say("  vs multi-pass", d)
        assert max(d) < 1e-12
    op64.set_algorithm(0)
    op64.set_groups(0)
    run_contig(op64, fs, device)
    assert op64.get_algorithm() == 7  # the FP64 default for isotropic scattering
    op32, _ = make_op(N, L, Nrho, M, Mpre, masses, 1.0, omegas, (1.0,) * (S * S), precision=32)
    op32.set_symmetry(sym)
    op32.set_algorithm(7)
    Q32 = run_contig(op32, fs, device)
    assert op32.get_algorithm() == 7
    e32 = [rel(Q32[p], Qm[p]) for p in range(S)]
    say("  fp32 slab vs fp64", e32)
    assert max(e32) < 1e-4


def test_slab_kernel_needs_isotropic_scattering(device):
    from dgfs_hip import _lib
    op, _ = make_op(16, 5.0, 4, 12, 24, (1.0, 2.1), 1.3, (0.8,) * 4, (1,) * 4, precision=64)
    with pytest.raises(_lib.DGFSUnsupported):
        op.set_algorithm(7)
    op8, _ = make_op(8, 5.0, 4, 12, 24, (1.0, 2.1), 1.0, (0.8,) * 4, (1,) * 4, precision=64)
    with pytest.raises(_lib.DGFSUnsupported):
        op8.set_algorithm(7)


def test_register_resident_needs_fp32_and_grid(device):
    from dgfs_hip import _lib
    op64, _ = make_op(16, 5.0, 4, 12, 24, (1.0, 2.1), 1.0, (0.8,) * 4, (1,) * 4, precision=64)
    with pytest.raises(_lib.DGFSUnsupported):
        op64.set_algorithm(4)
    op8, _ = make_op(8, 5.0, 4, 12, 24, (1.0, 2.1), 1.0, (0.8,) * 4, (1,) * 4, precision=32)
    with pytest.raises(_lib.DGFSUnsupported):
        op8.set_algorithm(4)
    with pytest.raises(_lib.DGFSInvalidValue):
        op8.set_algorithm(3)


def test_autotune_two_species(device, tmp_path):
    """The autotuner on the two-species operator: every candidate runs, the
    fastest is applied and cached, and the tuned operator gives the result of
    the default one (FP32, different summation orders: 1e-5)."""
    import torch
    from dgfs_hip.autotune import Autotuner
    N, L, Nrho, M, Mpre, masses = 16, 5.0, 4, 12, 24, (1.0, 2.1)
    op, _ = make_op(N, L, Nrho, M, Mpre, masses, 1.0, (0.81, 0.805, 0.805, 0.80), (1.0,) * 4, precision=32,
                    project=True)
    fs = _mixture_states(V.BiVelocityMesh(N, L, Nrho, M, masses), 32, seed=41)
    Q0 = run_contig(op, fs, device)
    ft = [torch.as_tensor(np.ascontiguousarray(f), dtype=torch.float64, device=device) for f in fs]
    Qt = [torch.empty_like(f) for f in ft]
    tuner = Autotuner(cache_path=str(tmp_path / "tune.json"), repeats=1)
    best, ms = tuner.tune(op, [f.data_ptr() for f in ft], [q.data_ptr() for q in Qt], 32)
    say("vss tune best", best, "of", len(ms))
    assert len(ms) == len(Autotuner.candidates(op, torch.cuda.get_device_properties(0).multi_processor_count))
    Q1 = run_contig(op, fs, device)
    assert max(rel(Q1[p], Q0[p]) for p in range(2)) < 1e-5
    best2, _ = Autotuner(cache_path=str(tmp_path / "tune.json")).tune(op, [f.data_ptr() for f in ft],
                                                                     [q.data_ptr() for q in Qt], 32)
    assert best2 == best
