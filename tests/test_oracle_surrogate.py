"""Pins of the surrogate checker (oracle/surrogate.py).  CPU only."""
import math

import numpy as np
import pytest

from oracle import surrogate as S


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    th = S.init_params(1, scale=0.7)
    P = 3
    x = rng.normal(size=(P, S.MODES)) * 0.3
    lt = rng.normal(size=P) * 0.2
    target = rng.normal(size=(P, S.MODES))

    def loss(t):
        out, _ = S.forward(t, x, lt)
        return 0.5 * np.sum((out - target) ** 2)

    out, cache = S.forward(th, x, lt)
    g = S.backward(th, cache, out - target)
    blocks = [(S.OFF_P0, S.WIDTH), (S.OFF_P1, S.WIDTH), (S.OFF_E, S.MODES * S.WIDTH)]
    for l in range(S.LAYERS):
        b = S.OFF_LAYERS + l * S.LAYER_SIZE
        blocks += [(b, S.MODES * S.MODES), (b + S.MODES * S.MODES, 2 * S.WIDTH * S.WIDTH + S.WIDTH)]
    blocks += [(S.OFF_POUT, S.WIDTH), (S.OFF_BOUT, S.MODES)]
    for off, size in blocks:
        for i in rng.choice(size, 3, replace=False):
            k = off + int(i)
            e = 1e-5
            tp, tm = th.copy(), th.copy()
            tp[k] += e
            tm[k] -= e
            fd = (loss(tp) - loss(tm)) / (2 * e)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (off, k, fd, g[k])


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0], dtype=np.float32)
    r = S.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0            # tie rounds to even
    assert r[2] == 1.0078125                       # above the tie rounds up
    assert abs(r[3] + 3.140625) < 1e-12
    assert r[4] == 0.0


@pytest.mark.parametrize("N,L", [(16, 6.0), (24, 7.0)])
def test_encoder_of_maxwellian(N, L):
    c0 = S.grid(N, L)
    I, J, K = np.meshgrid(c0, c0, c0, indexing="ij")
    rho, u, T = 1.3, np.array([0.3, -0.2, 0.1]), 1.1
    M = rho / (math.pi * T) ** 1.5 * np.exp(-((I - u[0]) ** 2 + (J - u[1]) ** 2 + (K - u[2]) ** 2) / T)
    mom, a, lt = S.encode(M.ravel(), N, L)
    # midpoint sums of a Gaussian converge spectrally in h: measured moment
    # errors 4e-8 at h = 0.75, 7e-13 at h = 0.58; the modes of a Maxwellian
    # (in the basis span up to those moment errors) follow them
    tol = 1e-7 if N == 16 else 2e-12
    assert abs(mom[0] - rho) < tol and np.max(np.abs(mom[1:4] - u)) < tol and abs(mom[4] - T) < tol
    assert mom[5] < tol and abs(lt - math.log(T)) < tol and mom[6] < 100 * tol
    assert np.max(np.abs(a)) < 100 * tol


@pytest.mark.parametrize("N,L,T", [(16, 6.0, 1.0), (32, 7.0, 1.2), (24, 7.0, 2.0)])
def test_projection_reproduces_the_basis_span(N, L, T):
    """targets(decode(y)) = y for any coefficients y: the projection onto the
    grid-orthonormal basis is exact on its span, on coarse and fine grids."""
    rng = np.random.default_rng(2)
    mom = np.array([0.9, 0.2, 0.1, -0.3, T, 0.0])
    y = rng.normal(size=S.MODES) * 0.1
    g = S.decode(y, mom, N, L, gamma=0.38, prefac=1.7, remove_invariants=False)
    back = S.targets(g, mom, N, L, 0.38, 1.7)
    # measured 3e-12 (h = 0.75) and 4e-13 (h = 0.44)
    assert np.max(np.abs(back - y)) < 1e-10


def test_decoder_removes_invariants():
    """Zeroing the mass / momentum modes and the trace of the second-order
    modes removes the invariants of the decoded field up to the grid
    quadrature (1e-8 at h = 0.58, 1e-16 at h = 0.44)."""
    N, L = 32, 7.0
    rng = np.random.default_rng(4)
    mom = np.array([1.1, -0.2, 0.15, 0.05, 0.9, 0.0])
    Q = S.decode(rng.normal(size=S.MODES), mom, N, L, gamma=0.38, prefac=1.0)
    c0 = S.grid(N, L)
    I, J, K = np.meshgrid(c0, c0, c0, indexing="ij")
    cw = (2 * L / N) ** 3
    sc = np.abs(Q).sum() * cw
    for phi in (np.ones_like(I), I, J, K, I ** 2 + J ** 2 + K ** 2):
        assert abs((phi.ravel() * Q).sum() * cw) / sc < 1e-12
