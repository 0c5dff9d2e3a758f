"""FP64 CPU oracle for the two-species fast-spectral VSS collision operator Q_pq.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: it is
the checker used by ``tests/``.  It restates, in numpy/scipy, the algorithm of
the reference's binary-mixture solver (file:line relative to the reference
repository root):

* velocity box, S and R from the mass ratio .. frfs/solvers/dgfsbi/velocitymesh.py:62-100
* radial rule on [0, R] ..................... frfs/solvers/dgfsbi/velocitymesh.py:126-127
* full-sphere design, sw = 4 pi / M ......... frfs/solvers/dgfsbi/velocitymesh.py:129-137
* Mpre design, swpre = 4 pi / Mpre .......... frfs/solvers/dgfsbi/scattering.py:51-56
* gamma, eta, prefactor per pair ............ frfs/solvers/dgfsbi/scattering.py:58-92
* tables a, b_pq (Nrho*M*n, complex), c_pq .. frfs/solvers/dgfsbi/scattering.py:172-197,
                                              kernels/scattering/vss-gll.mako:31-95
* Q_pq dataflow ............................. frfs/solvers/dgfsbi/scattering.py:259-320,
                                              vss-gll.mako:102-300
* species Maxwellian ........................ frfs/solvers/dgfsbi/initcond.py:66-73

How it is pinned (tests/test_oracle_vss.py):
  * brute force: the same discrete formula with explicit n x n DFT matrices
    at N = 4 (q_bruteforce);
  * reduction to the single-species VHS operator: equal masses, alpha = 1 and
    b, c built with the exact spherical average (sinc) instead of the Mpre
    design give the VHS oracle's Q(f, f) -- a different code path (cos/sin
    squares, half sphere, real tables) that is itself pinned by the BKW
    closed form;
  * equilibrium: Q_pq(M_p, M_q) ~ 0 for species Maxwellians with a common
    velocity and temperature (mass ratio 2.1);
  * conservation: species mass per pair, and m_p int c Q_pq + m_q int c Q_qp,
    m_p int |c|^2 Q_pq + m_q int |c|^2 Q_qp, measured against int |Q|.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .vhs import _fft3, _ifft3u, antipodal_design, gauss_jacobi, lvectors


def full_sphere_design(M):
    """Symmetric M-point design (2, 6, 12, 24 points); the reference loads the
    symmetric table with npts=M and keeps all rows (velocitymesh.py:129-137)."""
    if M % 2:
        raise ValueError("symmetric designs have an even number of points")
    return antipodal_design(M // 2)


@dataclass
class BiVelocityMesh:
    N: int
    L: float
    Nrho: int
    M: int
    masses: tuple
    sz: np.ndarray | None = field(default=None, repr=False)

    def __post_init__(self):
        N, L = self.N, self.L
        self.nspcs = len(self.masses)
        self.n = N ** 3
        self.K = self.M * self.Nrho
        mr = max(self.masses)
        # velocitymesh.py:72-76
        self.S = min(L * 2.0 / (max(4.0 * mr / (1.0 + mr), 2.0) + math.sqrt(1.0 + mr) + 1.0),
                     L * 2.0 / (max(4.0 / (1.0 + mr), 2.0) + math.sqrt(1.0 + 1.0 / mr) + 1.0))
        self.R = 2.0 * self.S
        self.cw = (2.0 * L / N) ** 3
        c0 = np.linspace(-L + L / N, L - L / N, N)
        I, J, Kk = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        self.cv = np.stack([c0[I.ravel()], c0[J.ravel()], c0[Kk.ravel()]])
        z, w = gauss_jacobi(self.Nrho)
        self.qz = (self.R / 2.0) * (1.0 + z)
        self.qw = (self.R / 2.0) * w
        if self.sz is None:
            self.sz = full_sphere_design(self.M)
        self.sw = 4.0 * math.pi / self.M


def vss_prefactor(mp, mq, omega, alpha, dRef, Tref, H0, T0, n0):
    """scattering.py:74-86."""
    invKn = math.sqrt(1.0 + mp / mq) * math.pi * n0 * dRef * dRef * (Tref / T0) ** (omega - 0.5) * H0
    mu = mp * mq / (mp + mq)
    return invKn * alpha / (math.sqrt(1.0 + mp / mq) * mu ** (omega - 0.5)
                            * 2.0 ** (1.0 + alpha) * math.gamma(2.5 - omega) * math.pi)


def vss_tables(vm: BiVelocityMesh, mp, mq, gamma, eta, szpre=None, exact_sphere=False):
    """b (Nrho*M, n) complex and c (n) complex of pair (p, q), vss-gll.mako:53-95.

    b_xy(l) = swpre qz_x^{gamma+2} sum_j (1 + s_y.s'_j)^eta exp(-i mpq a'_j(l)),
    a'_j = (pi/L) qz_x (s'_j . l), mpq = m_q / (m_p + m_q);
    c(l) = sum_xy qw_x sw swpre qz_x^{gamma+2} sum_j (1 + s_y.s'_j)^eta exp(-i a'_j(l)).

    exact_sphere=True (alpha = 1 only) replaces the Mpre-point sum by its
    limit, the spherical average of a plane wave, 4 pi sinc(k |l|): the form
    in which the operator reduces to the single-species VHS tables."""
    L, n = vm.L, vm.n
    l = lvectors(vm.N)
    mpq = mq / (mp + mq)
    b = np.empty((vm.Nrho, vm.M, n), dtype=np.complex128)
    c = np.zeros(n, dtype=np.complex128)
    if exact_sphere:
        if eta != 0.0:
            raise ValueError("the exact spherical average needs alpha = 1 (eta = 0)")
        lnorm = np.sqrt(np.sum(l * l, axis=0))
        sinc = lambda s: np.sin(s + 1e-15) / (s + 1e-15)
        for x in range(vm.Nrho):
            fac = math.pi / L * vm.qz[x]
            qg = vm.qz[x] ** (gamma + 2.0)
            bx = qg * 4.0 * math.pi * sinc(mpq * fac * lnorm)
            b[x, :, :] = bx[None, :]
            c += vm.qw[x] * vm.sw * vm.M * qg * 4.0 * math.pi * sinc(fac * lnorm)
        return b.reshape(vm.K, n), c
    szpre = np.asarray(szpre, dtype=float)
    swpre = 4.0 * math.pi / len(szpre)
    for x in range(vm.Nrho):
        fac = math.pi / L * vm.qz[x]
        fac_b = swpre * vm.qz[x] ** (gamma + 2.0)
        fac_c = vm.qw[x] * vm.sw * fac_b
        Ld = fac * (szpre @ l)                          # (Mpre, n)
        eb = np.exp(-1j * mpq * Ld)
        ec = np.exp(-1j * Ld)
        for y in range(vm.M):
            B = (1.0 + szpre @ vm.sz[y]) ** eta          # (Mpre,)
            b[x, y] = fac_b * (B @ eb)
            c += fac_c * (B @ ec)
    return b.reshape(vm.K, n), c


class VSSOracle:
    """Q_pq(f_p, f_q) for every ordered species pair (scattering.py:259-330)."""

    def __init__(self, vm: BiVelocityMesh, gamma, eta, prefac, szpre=None, exact_sphere=False):
        """gamma, eta, prefac: dicts keyed by (p, q)."""
        self.vm = vm
        self.gamma, self.eta, self.prefac = gamma, eta, prefac
        m = vm.masses
        l = lvectors(vm.N)
        self.a = np.stack([math.pi / vm.L * vm.qz[x] * (vm.sz[y] @ l)
                           for x in range(vm.Nrho) for y in range(vm.M)])   # (K, n)
        self.tab = {}
        for p in range(vm.nspcs):
            for q in range(vm.nspcs):
                self.tab[p, q] = vss_tables(vm, m[p], m[q], gamma[p, q], eta[p, q],
                                            szpre, exact_sphere)

    def qij(self, p, q, fp, fq, workers=1):
        vm = self.vm
        N, n, Nrho, M = vm.N, vm.n, vm.Nrho, vm.M
        sh3 = (N, N, N)
        mp, mq = vm.masses[p], vm.masses[q]
        b, c = self.tab[p, q]
        F = _fft3(fp.reshape(sh3).astype(np.complex128), workers).ravel() / n
        G = _fft3(fq.reshape(sh3).astype(np.complex128), workers).ravel() / n
        # cosSinMul (mako:136-157): t1 = e^{+i a mq/(mp+mq)} F, t2 = e^{-i a mp/(mp+mq)} G
        t1 = np.exp(1j * self.a * (mq / (mp + mq))) * F
        t2 = np.exp(-1j * self.a * (mp / (mp + mq))) * G
        t3 = _ifft3u(t1.reshape((vm.K,) + sh3), workers)
        t1 = _ifft3u(t2.reshape((vm.K,) + sh3), workers)
        s = t3 * t1                                        # prod (mako:163-188)
        t1 = _fft3(s, workers).reshape(vm.K, n)
        w = np.repeat(vm.qw * vm.sw / n, M)                # computeQG (mako:224-250)
        QG = np.einsum("k,kr,kr->r", w, t1, b)
        QG = _ifft3u(QG.reshape(sh3), workers).ravel()
        loss = _ifft3u((c * G).reshape(sh3), workers).ravel()   # ax2 + IFFT
        return self.prefac[p, q] * (QG.real - fp * loss.real)   # output (mako:275-300)

    def q(self, fs, workers=1):
        """Q_p = sum_q Q_pq(f_p, f_q) for every species (scattering.py:323-330).
        fs: list of (..., n) arrays, one per species."""
        fs = [np.asarray(f, dtype=np.float64) for f in fs]
        batch = fs[0].shape[:-1]
        fl = [f.reshape(-1, self.vm.n) for f in fs]
        out = [np.zeros_like(f) for f in fl]
        for i in range(fl[0].shape[0]):
            for p in range(self.vm.nspcs):
                for q in range(self.vm.nspcs):
                    out[p][i] += self.qij(p, q, fl[p][i], fl[q][i], workers)
        return [o.reshape(batch + (self.vm.n,)) for o in out]


def q_bruteforce(vm: BiVelocityMesh, orc: VSSOracle, p, q, fp, fq):
    """Q_pq with explicit n x n DFT matrices (tiny N)."""
    N, n = vm.N, vm.n
    x = np.arange(N)
    W1 = np.exp(-2j * np.pi * np.outer(x, x) / N)
    Wf = np.kron(np.kron(W1, W1), W1)
    Wi = np.conj(Wf)
    mp, mq = vm.masses[p], vm.masses[q]
    b, c = orc.tab[p, q]
    F = Wf @ fp.astype(np.complex128) / n
    G = Wf @ fq.astype(np.complex128) / n
    QG = np.zeros(n, dtype=np.complex128)
    for k in range(vm.K):
        xr = k // vm.M
        a = orc.a[k]
        u = Wi @ (np.exp(1j * a * mq / (mp + mq)) * F)
        v = Wi @ (np.exp(-1j * a * mp / (mp + mq)) * G)
        QG += (vm.qw[xr] * vm.sw / n) * b[k] * (Wf @ (u * v))
    gain = Wi @ QG
    loss = Wi @ (c * G)
    return orc.prefac[p, q] * (gain.real - fp * loss.real)


def species_maxwellian(vm: BiVelocityMesh, p, nden, u, T):
    """initcond.py:66-73: n_p (m_p / (pi T))^{3/2} exp(-m_p |c - u|^2 / T)."""
    m = vm.masses[p]
    u = np.asarray(u, dtype=float).reshape(3, 1)
    return nden * (m / (math.pi * T)) ** 1.5 * np.exp(-m * np.sum((vm.cv - u) ** 2, axis=0) / T)
