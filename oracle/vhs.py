"""FP64 CPU oracle for the DGFS fast-spectral VHS collision operator Q(f, f).

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module: it is
the checker used by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``.  It restates, in numpy/scipy, the
algorithm of the reference solver (file:line citations are relative to the
reference repository root):

* velocity grid, weights and radii ........ frfs/solvers/dgfs/velocitymesh.py:62-84
* radial Gauss-Jacobi rule ................ frfs/quadratures.py:6-55,
                                            frfs/solvers/dgfs/velocitymesh.py:112-126
* half-sphere design, sw = 2*pi/M ......... frfs/solvers/dgfs/velocitymesh.py:128-135
* frequency vectors l ..................... frfs/solvers/dgfs/scattering.py:79-89
* tables aa, b1, b2 ....................... frfs/solvers/dgfs/kernels/scattering/vhs-gll.mako:11-14,45-101
* prefactor ............................... frfs/solvers/dgfs/scattering.py:50-62
* Q(f) dataflow (6 FFTs + 6 kernels) ...... frfs/solvers/dgfs/scattering.py:227-286,
                                            vhs-gll.mako:103-266
* Maxwellian / BKW initial states ......... frfs/solvers/dgfs/initcond.py:55-58,130-137

How each function is pinned (tests/test_oracle_*.py):
  gauss_jacobi / gauss_lobatto_jacobi -- numpy.polynomial.legendre.leggauss,
      scipy.special.roots_jacobi and the reference's tabulated line rules
      (frfs/quadrules/line/gauss-legendre-n3-d5-spu.txt etc.) as data.
  half_sphere_design -- the reference's tabulated symmetric designs (M=3, 6)
      as data; M=12 has no reference table and is checked as an antipodal
      spherical 5-design from the mathematics.
  vhs_tables -- closed forms (b2 at l=0, aa linear in l, sinc limits).
  q_spectral -- brute-force DFT evaluation (q_bruteforce, explicit n x n
      matrices) at N=4 and 6; the BKW closed-form time derivative for Maxwell
      molecules; Q(Maxwellian) ~ 0; translation equivariance; conservation.
  vhs_prefactor -- the VHS kernel derived from Bird's cross-section in SI
      units and the hard-sphere mean free path (tests/test_prefactor.py).
  moments5 -- Maxwellian and BKW closed-form moments (tests/test_oracle_bkw.py).
  project_invariants -- constraints, idempotence, correction in the row space
      of the constraints (tests/test_oracle_vhs.py).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

try:
    import scipy.fft as _sfft
except Exception:  # pragma: no cover - scipy is present in this image
    _sfft = None


# ---------------------------------------------------------------------------
# Quadrature (Golub-Welsch), restating frfs/quadratures.py:6-55
# ---------------------------------------------------------------------------
def _jacobi_recurrence(n, a, b):
    """Monic Jacobi three-term recurrence coefficients (alpha_k, beta_k)."""
    alpha = np.zeros(n)
    beta = np.zeros(n)
    for k in range(n):
        s = 2.0 * k + a + b
        if k == 0:
            alpha[k] = (b - a) / (a + b + 2.0)
        else:
            alpha[k] = (b * b - a * a) / (s * (s + 2.0))
    beta[0] = 2.0 ** (a + b + 1.0) * math.gamma(a + 1.0) * math.gamma(b + 1.0) \
        / math.gamma(a + b + 2.0)
    for k in range(1, n):
        s = 2.0 * k + a + b
        num = 4.0 * k * (k + a) * (k + b) * (k + a + b)
        den = s * s * (s + 1.0) * (s - 1.0)
        beta[k] = num / den
    return alpha, beta


def _golub_welsch(alpha, beta):
    J = np.diag(alpha) + np.diag(np.sqrt(beta[1:]), 1) + np.diag(np.sqrt(beta[1:]), -1)
    z, V = np.linalg.eigh(J)
    idx = np.argsort(z)
    z = z[idx]
    w = beta[0] * V[0, idx] ** 2
    return z, w


def gauss_jacobi(n, a=0.0, b=0.0):
    """n-point Gauss-Jacobi rule on [-1, 1] (frfs/quadratures.py:40-43)."""
    alpha, beta = _jacobi_recurrence(n, a, b)
    return _golub_welsch(alpha, beta)


def gauss_lobatto_jacobi(n, a=0.0, b=0.0):
    """n-point Gauss-Lobatto-Jacobi rule (frfs/quadratures.py:45-55).

    The last recurrence pair is modified so that the monic degree-n
    polynomial vanishes at -1 and +1 (Golub 1973)."""
    if n < 2:
        raise ValueError("Lobatto rule needs n >= 2")
    alpha, beta = _jacobi_recurrence(n, a, b)

    def monic(x):
        p0, p1 = 1.0, x - alpha[0]
        if n - 1 == 0:
            return 0.0, p0
        for k in range(1, n - 1):
            p0, p1 = p1, (x - alpha[k]) * p1 - beta[k] * p0
        return p0, p1  # p_{n-2}(x), p_{n-1}(x)

    pm2_m, pm1_m = monic(-1.0)
    pm2_p, pm1_p = monic(1.0)
    # (x - alpha') p_{n-1} - beta' p_{n-2} = 0 at x = -1, +1
    A = np.array([[pm1_m, pm2_m], [pm1_p, pm2_p]])
    rhs = np.array([-1.0 * pm1_m, 1.0 * pm1_p])
    al, be = np.linalg.solve(A, rhs)
    alpha = alpha.copy()
    beta = beta.copy()
    alpha[n - 1] = al
    beta[n - 1] = be
    return _golub_welsch(alpha, beta)


# ---------------------------------------------------------------------------
# Spherical designs (half-sphere), frfs/sphericaldesign/__init__.py:59-107
# and frfs/solvers/dgfs/velocitymesh.py:128-135 (first M rows of 2M points)
# ---------------------------------------------------------------------------
def antipodal_design(M):
    """2M-point antipodal design whose first M rows are the half-sphere set.

    M=3: the octahedron (reference table womersley-ss003-m00006.txt).
    M=6: the icosahedron in the reference's ordering (womersley-ss005-m00012).
    M=12: the compound of two icosahedra, i.e. every permutation of
          (0, +-a, +-b) with a^4 + b^4 = 3/5 -- an antipodal 5-design with 24
          points.  The reference ships no 24-point table."""
    if M == 3:
        half = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    elif M == 6:
        zc = 1.0 / math.sqrt(5.0)
        rc = 2.0 / math.sqrt(5.0)
        az = [0.0, -72.0, 144.0, -144.0, 72.0]
        half = [[0.0, 0.0, 1.0]]
        for d in az:
            t = math.radians(d)
            half.append([rc * math.cos(t), rc * math.sin(t), zc])
        half = np.array(half)
    elif M == 12:
        s = 0.5 * (1.0 + 1.0 / math.sqrt(5.0))
        a, b = math.sqrt(s), math.sqrt(1.0 - s)
        half = []
        for u, v in ((a, b), (b, a)):
            for sv in (1.0, -1.0):
                half.append([0.0, u, sv * v])
                half.append([u, 0.0, sv * v])
                half.append([u, sv * v, 0.0])
        half = np.array(half)
        # order: group by pattern for readability
        order = np.lexsort((half[:, 2], half[:, 1], half[:, 0]))[::-1]
        half = half[order]
    else:
        raise ValueError(f"no spherical design with {2*M} points")
    return np.vstack([half, -half])


def half_sphere_design(M):
    return antipodal_design(M)[:M].copy()


# ---------------------------------------------------------------------------
# Velocity mesh, frfs/solvers/dgfs/velocitymesh.py:7-161
# ---------------------------------------------------------------------------
@dataclass
class VelocityMesh:
    N: int
    L: float
    Nrho: int
    M: int
    quad_rule: str = "jacobi"
    quad_a: float = 0.0
    quad_b: float = 0.0
    sz: np.ndarray | None = field(default=None, repr=False)

    def __post_init__(self):
        N, L = self.N, self.L
        self.n = N ** 3
        self.K = self.M * self.Nrho
        self.S = L * 2.0 / (3.0 + math.sqrt(2.0))          # velocitymesh.py:63
        self.R = 2.0 * self.S                                # velocitymesh.py:64
        self.cw = (2.0 * L / N) ** 3                         # velocitymesh.py:72
        c0 = np.linspace(-L + L / N, L - L / N, N)           # velocitymesh.py:73-74
        I, J, Kk = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
        # flat index r = I*N^2 + J*N + K (velocitymesh.py:77-84)
        self.cv = np.stack([c0[I.ravel()], c0[J.ravel()], c0[Kk.ravel()]])
        if self.quad_rule == "jacobi":
            z, w = gauss_jacobi(self.Nrho, self.quad_a, self.quad_b)
        elif self.quad_rule == "lobatto":
            z, w = gauss_lobatto_jacobi(self.Nrho, self.quad_a, self.quad_b)
        else:
            raise ValueError("quad-rule must be jacobi or lobatto")
        self.qz = (self.R / 2.0) * (1.0 + z)                 # velocitymesh.py:125
        self.qw = (self.R / 2.0) * w                         # velocitymesh.py:126
        if self.sz is None:
            self.sz = half_sphere_design(self.M)
        self.sw = 2.0 * math.pi / self.M                     # velocitymesh.py:135


def lvectors(N):
    """Integer frequency vectors in FFT order (scattering.py:79-89)."""
    l0 = np.concatenate((np.arange(0, N // 2), np.arange(-N // 2, 0))).astype(float)
    I, J, K = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return np.stack([l0[I.ravel()], l0[J.ravel()], l0[K.ravel()]])


def _sinc(s):
    # vhs-gll.mako:11-14: sin(s + 1e-15)/(s + 1e-15)
    return np.sin(s + 1e-15) / (s + 1e-15)


def vhs_tables(vm: VelocityMesh, gamma: float):
    """aa (K, n), b1 (Nrho, n), b2 (n): vhs-gll.mako:45-101."""
    L = vm.L
    l = lvectors(vm.N)
    lnorm = np.sqrt(np.sum(l * l, axis=0))
    aa = np.empty((vm.Nrho, vm.M, vm.n))
    b1 = np.empty((vm.Nrho, vm.n))
    b2 = np.zeros(vm.n)
    for p in range(vm.Nrho):
        fac = math.pi / L * vm.qz[p] / 2.0
        for q in range(vm.M):
            aa[p, q] = fac * (vm.sz[q] @ l)
        b1[p] = vm.qz[p] ** (gamma + 2.0) * 4.0 * math.pi * _sinc(fac * lnorm)
        b2 += vm.qw[p] * vm.qz[p] ** (gamma + 2.0) * 16.0 * math.pi ** 2 \
            * _sinc(math.pi / L * vm.qz[p] * lnorm)
    return aa.reshape(vm.K, vm.n), b1, b2


def vhs_prefactor(omega, dRef, Tref, H0, T0, n0, alpha=1.0):
    """scattering.py:50-62."""
    invKn = H0 * math.sqrt(2.0) * math.pi * n0 * dRef * dRef * (Tref / T0) ** (omega - 0.5)
    return invKn * alpha / (2.0 ** (2.0 - omega + alpha) * math.gamma(2.5 - omega) * math.pi)


# ---------------------------------------------------------------------------
# FFT helpers: unnormalised forward (e^-) and inverse (e^+), row-major,
# last index fastest -- the cuFFT Z2Z convention (cufft.py:182-183).
# ---------------------------------------------------------------------------
def _fft3(x, workers):
    if _sfft is not None:
        return _sfft.fftn(x, axes=(-3, -2, -1), workers=workers)
    return np.fft.fftn(x, axes=(-3, -2, -1))


def _ifft3u(x, workers):
    n = x.shape[-1] * x.shape[-2] * x.shape[-3]
    if _sfft is not None:
        return _sfft.ifftn(x, axes=(-3, -2, -1), workers=workers) * n
    return np.fft.ifftn(x, axes=(-3, -2, -1)) * n


# ---------------------------------------------------------------------------
# Q(f): Appendix-A dataflow of scattering.py:227-286
# ---------------------------------------------------------------------------
class VHSOracle:
    def __init__(self, vm: VelocityMesh, gamma: float, prefac: float):
        self.vm = vm
        self.gamma = gamma
        self.prefac = prefac
        self.aa, self.b1, self.b2 = vhs_tables(vm, gamma)
        self.cos_a = np.cos(self.aa)
        self.sin_a = np.sin(self.aa)

    def q(self, f, workers=1):
        """Q for one point or a batch: f has shape (..., n) -> (..., n)."""
        f = np.asarray(f, dtype=np.float64)
        batch = f.shape[:-1]
        fl = f.reshape(-1, self.vm.n)
        out = np.empty_like(fl)
        for i in range(fl.shape[0]):
            out[i] = self._q_one(fl[i], workers)
        return out.reshape(batch + (self.vm.n,))

    def _q_one(self, f, workers):
        vm = self.vm
        N, n, Nrho, M = vm.N, vm.n, vm.Nrho, vm.M
        sh3 = (N, N, N)
        # r2z + forward FFT, then cosSinMul's 1/n scaling (mako:120)
        FTf = _fft3(f.reshape(sh3).astype(np.complex128), workers).ravel() / n
        # t1 = cos(a) FTf, t2 = sin(a) FTf  (mako:124-138)
        t1 = (self.cos_a * FTf).reshape((vm.K,) + sh3)
        t2 = (self.sin_a * FTf).reshape((vm.K,) + sh3)
        t3 = _ifft3u(t1, workers)
        t4 = _ifft3u(t2, workers)
        # magSqr: complex squares, not moduli (mako:160-164)
        s = t3 * t3 + t4 * t4
        t1 = _fft3(s, workers).reshape(Nrho, M, n)
        # computeQG (mako:192-205)
        w = 2.0 * vm.qw * vm.sw / n
        QG = np.einsum("p,pr,pr->r", w, self.b1, t1.sum(axis=1))
        QG = _ifft3u(QG.reshape(sh3), workers).ravel()
        # ax + inverse FFT (mako:209-220; scattering.py:276-281)
        fC = _ifft3u((self.b2 * FTf).reshape(sh3), workers).ravel()
        # output (mako:264)
        return self.prefac * (QG.real - f * fC.real)


def q_bruteforce(f, vm: VelocityMesh, gamma, prefac):
    """The same discrete formula with explicit n x n DFT matrices (tiny N)."""
    N, n = vm.N, vm.n
    x = np.arange(N)
    W1 = np.exp(-2j * np.pi * np.outer(x, x) / N)     # forward, e^-
    Wf = np.kron(np.kron(W1, W1), W1)                  # row-major 3D DFT
    Wi = np.conj(Wf)                                   # unnormalised inverse
    aa, b1, b2 = vhs_tables(vm, gamma)
    F = (Wf @ f.astype(np.complex128)) / n
    QG = np.zeros(n, dtype=np.complex128)
    for p in range(vm.Nrho):
        acc = np.zeros(n, dtype=np.complex128)
        for q in range(vm.M):
            a = aa[p * vm.M + q]
            t3 = Wi @ (np.cos(a) * F)
            t4 = Wi @ (np.sin(a) * F)
            acc += Wf @ (t3 * t3 + t4 * t4)
        QG += (2.0 * vm.qw[p] * vm.sw / n) * b1[p] * acc
    G = Wi @ QG
    Lf = Wi @ (b2 * F)
    return prefac * (G.real - f * Lf.real)


# ---------------------------------------------------------------------------
# States, initcond.py:55-58 and 130-137
# ---------------------------------------------------------------------------
def maxwellian(vm: VelocityMesh, rho, u, T):
    u = np.asarray(u, dtype=float).reshape(3, 1)
    return rho / (math.pi * T) ** 1.5 * np.exp(-np.sum((vm.cv - u) ** 2, axis=0) / T)


def bkw(vm: VelocityMesh, t0):
    K = 1.0 - math.exp(-t0 / 6.0)
    c2 = np.sum(vm.cv * vm.cv, axis=0)
    return 1.0 / (2.0 * (2.0 * math.pi * K) ** 1.5) * np.exp(-c2 / (2.0 * K)) \
        * ((5.0 * K - 3.0) / K + (1.0 - K) / (K * K) * c2)


def bkw_dfdt(vm: VelocityMesh, t0):
    """Closed-form time derivative of the BKW solution, dK/dt = exp(-t/6)/6."""
    K = 1.0 - math.exp(-t0 / 6.0)
    dK = math.exp(-t0 / 6.0) / 6.0
    c2 = np.sum(vm.cv * vm.cv, axis=0)
    pre = 1.0 / (2.0 * (2.0 * math.pi * K) ** 1.5) * np.exp(-c2 / (2.0 * K))
    poly = (5.0 * K - 3.0) / K + (1.0 - K) / (K * K) * c2
    dpre = pre * (-1.5 / K + c2 / (2.0 * K * K)) * dK
    dpoly = (3.0 / (K * K) + c2 * (K - 2.0) / K ** 3) * dK
    return dpre * poly + pre * dpoly


def moments5(vm: VelocityMesh, f):
    """(rho, Ux, Uy, Uz, T) with T = (2/3) sum f |c-U|^2 cw / rho
    (initcond.py:70-92)."""
    f = np.asarray(f)
    rho = f.sum(axis=-1) * vm.cw
    U = np.stack([(f * vm.cv[d]).sum(axis=-1) * vm.cw / rho for d in range(3)], axis=-1)
    c2 = np.zeros_like(f)
    for d in range(3):
        c2 = c2 + (vm.cv[d] - U[..., d:d + 1]) ** 2
    T = (f * c2).sum(axis=-1) * vm.cw * (2.0 / 3.0) / rho
    return rho, U, T


def collision_invariants(vm: VelocityMesh, f):
    """Integrals of f against {1, cx, cy, cz, |c|^2} times cw."""
    f = np.asarray(f)
    c2 = np.sum(vm.cv ** 2, axis=0)
    return np.stack([f.sum(-1), (f * vm.cv[0]).sum(-1), (f * vm.cv[1]).sum(-1),
                     (f * vm.cv[2]).sum(-1), (f * c2).sum(-1)], axis=-1) * vm.cw


def project_invariants(vm: VelocityMesh, Q):
    """Orthogonal projection of Q (..., n) onto {Q : sum_r Q_r phi_j(c_r) cw = 0,
    j = 1..5} with phi = (1, cx, cy, cz, |c|^2): the minimum-norm correction
    Q - C^+ C Q, C the 5 x n constraint matrix (pseudo-inverse by SVD).
    This is what the FP64 projection of the HIP operators computes (by the
    normal equations).  Pinned by its defining properties: the constraints hold,
    the map is idempotent, the correction lies in the row space of C (so no
    other point of the constraint set is closer)."""
    Q = np.asarray(Q, dtype=np.float64)
    c = vm.cv
    C = np.stack([np.ones(vm.n), c[0], c[1], c[2], (c * c).sum(0)]) * vm.cw
    flat = Q.reshape(-1, vm.n)
    out = flat - (np.linalg.pinv(C) @ (C @ flat.T)).T
    return out.reshape(Q.shape)
