"""FP64 numpy restatement of the learned collision surrogate (checker only).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP surrogate kernels
(dgfs-be-solver_amd/csrc/surrogate.hip); nothing in the product imports it.
The surrogate has no counterpart in the reference solver (the reference only
has the spectral operator, frfs/solvers/dgfs/scattering.py:227-286); what is
restated here is the surrogate's own definition:

* Hermite encoder: moments rho, u, T of f (DGFS Maxwellian convention of
  frfs/solvers/dgfs/initcond.py:55-58, per-axis variance T/2), indicators
  ||f - M[f]|| / ||f|| and the share of the modes of total degree >= 4, and the coefficients a_n of f / C (C the Maxwellian's
  prefactor) in the grid-orthonormal basis spanned by M psi_n(c'),
  psi_n = prod He_{n_d}/sqrt(n_d!), n_d < 8, c' = (c - u)/sqrt(T/2) (the
  Maxwellian's own coefficient subtracted);
* targets b^ = (the same coefficients of Q) / (prefac rho T^{gamma/2});
* decoder Q = M[f] sum_n b_n psi_n(c') with the mass, momentum and energy
  modes removed;
* the 4-layer spectral network and its gradient (reverse mode written out by
  hand).

Pins (tests/test_oracle_surrogate.py): the backward pass against central
finite differences of the forward pass; the encoder against exact moments of
Maxwellians; the projection reproduces any field of the basis span
(coefficients of decode(y) are y); the invariant removal against the discrete
moments of the decoded field.
"""
from __future__ import annotations

import math

import numpy as np
from numpy.polynomial.hermite_e import hermevander

D = 8
MODES = D ** 3
WIDTH = 32
LAYERS = 4
LAYER_SIZE = MODES * MODES + 2 * WIDTH * WIDTH + WIDTH
OFF_P0, OFF_P1, OFF_E = 0, WIDTH, 2 * WIDTH
OFF_LAYERS = OFF_E + MODES * WIDTH
OFF_POUT = OFF_LAYERS + LAYERS * LAYER_SIZE
OFF_BOUT = OFF_POUT + WIDTH
NPARAMS = OFF_BOUT + MODES


def unpack(th):
    th = np.asarray(th, dtype=np.float64)
    P0 = th[OFF_P0:OFF_P0 + WIDTH]
    P1 = th[OFF_P1:OFF_P1 + WIDTH]
    E = th[OFF_E:OFF_E + MODES * WIDTH].reshape(MODES, WIDTH)
    layers = []
    for l in range(LAYERS):
        b = OFF_LAYERS + l * LAYER_SIZE
        K = th[b:b + MODES * MODES].reshape(MODES, MODES)
        b += MODES * MODES
        W = th[b:b + WIDTH * WIDTH].reshape(WIDTH, WIDTH)
        b += WIDTH * WIDTH
        R = th[b:b + WIDTH * WIDTH].reshape(WIDTH, WIDTH)
        b += WIDTH * WIDTH
        bias = th[b:b + WIDTH]
        layers.append((K, W, R, bias))
    Pout = th[OFF_POUT:OFF_POUT + WIDTH]
    bout = th[OFF_BOUT:OFF_BOUT + MODES]
    return P0, P1, E, layers, Pout, bout


def init_params(seed=0, scale=1.0):
    """A seeded initialisation (the one the host code uses as well)."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAMS)
    th[OFF_P0:OFF_P0 + WIDTH] = rng.normal(size=WIDTH) * scale
    th[OFF_P1:OFF_P1 + WIDTH] = rng.normal(size=WIDTH) * 0.1 * scale
    th[OFF_E:OFF_E + MODES * WIDTH] = rng.normal(size=MODES * WIDTH) * 0.1 * scale
    for l in range(LAYERS):
        b = OFF_LAYERS + l * LAYER_SIZE
        th[b:b + MODES * MODES] = rng.normal(size=MODES * MODES) / math.sqrt(MODES) * scale
        b += MODES * MODES
        th[b:b + WIDTH * WIDTH] = rng.normal(size=WIDTH * WIDTH) / math.sqrt(WIDTH) * scale
        b += WIDTH * WIDTH
        th[b:b + WIDTH * WIDTH] = rng.normal(size=WIDTH * WIDTH) / math.sqrt(WIDTH) * scale
    th[OFF_POUT:OFF_POUT + WIDTH] = rng.normal(size=WIDTH) / math.sqrt(WIDTH) * scale
    return th


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _gelu(x):
    k0, k1 = math.sqrt(2.0 / math.pi), 0.044715
    return 0.5 * x * (1.0 + np.tanh(k0 * (x + k1 * x ** 3)))


def _gelu_d(x):
    k0, k1 = math.sqrt(2.0 / math.pi), 0.044715
    t = np.tanh(k0 * (x + k1 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * k0 * (1 + 3 * k1 * x * x)


def forward(th, x, lt, round_bf16=False):
    """x [P][512], lt [P] -> out [P][512]; activations Z[m][p][c]."""
    r = bf16 if round_bf16 else (lambda a: a)
    P0, P1, E, layers, Pout, bout = unpack(th)
    x = np.asarray(x, dtype=np.float64)
    lt = np.asarray(lt, dtype=np.float64)
    Z = P0[None, None, :] * x.T[:, :, None] + P1[None, None, :] * lt[None, :, None] + E[:, None, :]
    cache = {"x": x, "lt": lt, "Z": [Z], "U": [], "V": []}
    for l, (K, W, R, bias) in enumerate(layers):
        U = np.einsum("mk,kpc->mpc", r(K), r(Z))
        V = np.einsum("mpc,cd->mpd", r(U), r(W)) + np.einsum("mpc,cd->mpd", r(Z), r(R)) + bias
        Z = _gelu(V) if l + 1 < LAYERS else V
        cache["U"].append(U)
        cache["V"].append(V)
        cache["Z"].append(Z)
    out = np.einsum("mpc,c->pm", Z, Pout) + bout[None, :]
    return out, cache


def backward(th, cache, dout):
    P0, P1, E, layers, Pout, bout = unpack(th)
    g = np.zeros(NPARAMS)
    Z4 = cache["Z"][LAYERS]
    g[OFF_POUT:OFF_POUT + WIDTH] = np.einsum("mpc,pm->c", Z4, dout)
    g[OFF_BOUT:OFF_BOUT + MODES] = dout.sum(0)
    dZ = dout.T[:, :, None] * Pout[None, None, :]
    for l in range(LAYERS - 1, -1, -1):
        K, W, R, bias = layers[l]
        V, U, Z = cache["V"][l], cache["U"][l], cache["Z"][l]
        dV = dZ * _gelu_d(V) if l + 1 < LAYERS else dZ
        b = OFF_LAYERS + l * LAYER_SIZE
        dU = np.einsum("mpd,cd->mpc", dV, W)
        g[b:b + MODES * MODES] = np.einsum("mpc,kpc->mk", dU, Z).ravel()
        b += MODES * MODES
        g[b:b + WIDTH * WIDTH] = np.einsum("mpc,mpd->cd", U, dV).ravel()
        b += WIDTH * WIDTH
        g[b:b + WIDTH * WIDTH] = np.einsum("mpc,mpd->cd", Z, dV).ravel()
        b += WIDTH * WIDTH
        g[b:b + WIDTH] = dV.sum((0, 1))
        dZ = np.einsum("mpd,cd->mpc", dV, R) + np.einsum("mk,mpc->kpc", K, dU)
    g[OFF_P0:OFF_P0 + WIDTH] = np.einsum("mpc,pm->c", dZ, cache["x"])
    g[OFF_P1:OFF_P1 + WIDTH] = np.einsum("mpc,p->c", dZ, cache["lt"])
    g[OFF_E:OFF_E + MODES * WIDTH] = dZ.sum(1).ravel()
    return g


# ---------------------------------------------------------------------------
# Hermite encoder / decoder
# ---------------------------------------------------------------------------
def grid(N, L):
    h = 2.0 * L / N
    return -L + 0.5 * h + h * np.arange(N)


def moments(f, N, L):
    c0 = grid(N, L)
    cw = (2.0 * L / N) ** 3
    F = f.reshape(N, N, N)
    rho = F.sum() * cw
    u = np.array([np.einsum("ijk,i->", F, c0), np.einsum("ijk,j->", F, c0), np.einsum("ijk,k->", F, c0)]) * cw / rho
    I, J, K = np.meshgrid(c0, c0, c0, indexing="ij")
    c2 = I ** 2 + J ** 2 + K ** 2
    T = (2.0 / 3.0) * ((F * c2).sum() * cw / rho - u @ u)
    M = rho / (math.pi * T) ** 1.5 * np.exp(-((I - u[0]) ** 2 + (J - u[1]) ** 2 + (K - u[2]) ** 2) / T)
    ind = np.linalg.norm(F - M) / np.linalg.norm(F)
    return rho, u, T, ind, M.ravel()


SQRT_FACT = np.sqrt([math.factorial(a) for a in range(D)])


def _vander(N, L, u, T):
    """Per-axis tables He_a(x)/sqrt(a!) exp(-x^2/2), x = (c - u)/sqrt(T/2):
    the 1D factors of the basis M psi_n, psi_n = prod_d He_{n_d}/sqrt(n_d!)."""
    c0 = grid(N, L)
    s = math.sqrt(0.5 * T)
    out = []
    for d in range(3):
        x = (c0 - u[d]) / s
        out.append(hermevander(x, D - 1) / SQRT_FACT[None, :] * np.exp(-0.5 * x * x)[:, None])
    return out


def _linv(V, hs):
    """Inverse Cholesky factors of the per-axis Gram matrices (h/s) V^T V."""
    return [np.linalg.inv(np.linalg.cholesky(hs * v.T @ v)) for v in V]


def hermite_project(g, N, L, u, T, C, scale):
    """Coefficients of g / (C scale) in the grid-orthonormal basis
    phi = psi L^-T (per axis; measure dx' = h/s), shape (8, 8, 8)."""
    V = _vander(N, L, u, T)
    hs = (2.0 * L / N) / math.sqrt(0.5 * T)
    Li = _linv(V, hs)
    B = np.einsum("ijk,ia,jb,kc->abc", g.reshape(N, N, N), V[0], V[1], V[2]) * hs ** 3
    return np.einsum("pa,qb,rc,abc->pqr", Li[0], Li[1], Li[2], B) / (C * scale)


def encode(f, N, L):
    rho, u, T, ind, _ = moments(f, N, L)
    C = rho / (math.pi * T) ** 1.5
    a = hermite_project(f, N, L, u, T, C, 1.0).ravel()
    deg = np.add.outer(np.add.outer(np.arange(D), np.arange(D)), np.arange(D)).ravel()
    ind_hi = math.sqrt(np.sum(a[deg >= 4] ** 2) / np.sum(a ** 2))
    V = _vander(N, L, u, T)
    Li = _linv(V, (2.0 * L / N) / math.sqrt(0.5 * T))
    a[0] -= 1.0 / (Li[0][0, 0] * Li[1][0, 0] * Li[2][0, 0])   # the Maxwellian's own coefficient
    return np.array([rho, *u, T, ind, ind_hi]), a, math.log(T)


def targets(Q, mom, N, L, gamma, prefac):
    rho, u, T = mom[0], mom[1:4], mom[4]
    C = rho / (math.pi * T) ** 1.5
    return hermite_project(Q, N, L, u, T, C, prefac * rho * T ** (0.5 * gamma)).ravel()


def decode(y, mom, N, L, gamma, prefac, remove_invariants=True):
    rho, u, T = mom[0], mom[1:4], mom[4]
    V = _vander(N, L, u, T)
    Li = _linv(V, (2.0 * L / N) / math.sqrt(0.5 * T))
    c = (prefac * rho * T ** (0.5 * gamma)) * np.asarray(y, dtype=np.float64).reshape(D, D, D)
    b = np.einsum("ap,bq,cr,abc->pqr", Li[0], Li[1], Li[2], c)   # L^-T per axis
    if remove_invariants:
        tr = (b[2, 0, 0] + b[0, 2, 0] + b[0, 0, 2]) / 3.0
        b[2, 0, 0] -= tr
        b[0, 2, 0] -= tr
        b[0, 0, 2] -= tr
        b[0, 0, 0] = b[1, 0, 0] = b[0, 1, 0] = b[0, 0, 1] = 0.0
    C = rho / (math.pi * T) ** 1.5
    return (C * np.einsum("abc,ia,jb,kc->ijk", b, V[0], V[1], V[2])).ravel()


def project(Q, N, L):
    """Minimum-norm projection onto {1, c, |c|^2} (the product's FP64 projection)."""
    c0 = grid(N, L)
    cw = (2.0 * L / N) ** 3
    I, J, K = np.meshgrid(c0, c0, c0, indexing="ij")
    phi = np.stack([np.ones(N ** 3), I.ravel(), J.ravel(), K.ravel(), (I ** 2 + J ** 2 + K ** 2).ravel()]) * cw
    A = phi @ phi.T
    return Q - phi.T @ np.linalg.solve(A, phi @ Q)
