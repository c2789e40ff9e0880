"""Pins the CPU oracle of the warm-start network (oracle/fno_oracle.py):
gradients against central finite differences, the spectral convolution
against the identity on band-limited fields, the truncated DFT against a
brute-force sum, BF16 rounding and the first Adam step against their closed
forms."""
import numpy as np
import pytest

import fno_oracle as F

DIMS = (2, 4, 3, 2, 5)          # cin, C, M, L, P
N = 16


def _problem(seed=0, B=2):
    rng = np.random.default_rng(seed)
    _, n = F.layout(*DIMS)
    p = rng.normal(0, 0.4, n)
    # spectral weights large enough that the spectral path matters
    for name, off, shape in F.layout(*DIMS)[0]:
        if name.startswith("R"):
            p[off:off + int(np.prod(shape))] = rng.normal(0, 0.5 * N * N / 8, int(np.prod(shape))) / (N * N) * 8
    x = rng.normal(0, 1, (B, DIMS[0], N, N))
    tgt = rng.normal(0, 1, (B, 2, N, N))
    return p, x, tgt


def test_gradient_matches_finite_differences():
    p, x, tgt = _problem()
    loss, g = F.loss_grad(p, x, tgt, DIMS)
    rng = np.random.default_rng(1)
    lay, n = F.layout(*DIMS)
    # every parameter block: a few coordinates each
    for name, off, shape in lay:
        size = int(np.prod(shape))
        for idx in rng.choice(size, min(size, 4), replace=False):
            i = off + idx
            h = 1e-6 * max(1.0, abs(p[i]))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (F.loss_grad(pp, x, tgt, DIMS)[0] - F.loss_grad(pm, x, tgt, DIMS)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (name, idx, fd, g[i])
    # a random direction through all parameters at once
    v = rng.normal(0, 1, n)
    h = 1e-6
    fd = (F.loss_grad(p + h * v, x, tgt, DIMS)[0] - F.loss_grad(p - h * v, x, tgt, DIMS)[0]) / (2 * h)
    assert abs(fd - g @ v) <= 1e-6 * abs(fd)


def test_planted_sign_error_is_caught():
    """The check is sharp enough to see a conjugation mistake in dR: the
    imaginary part of a k != 0 spectral weight has a nonzero gradient and a
    flipped sign misses the finite difference by its full size."""
    p, x, tgt = _problem(3)
    _, g = F.loss_grad(p, x, tgt, DIMS)
    name, off, shape = [b for b in F.layout(*DIMS)[0] if b[0] == "R0"][0]
    M = DIMS[2]
    i = off + ((1 * 2 * M) + 1) * 2 + 1   # Im R[0][0][kx=1][ky=1]
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (F.loss_grad(pp, x, tgt, DIMS)[0] - F.loss_grad(pm, x, tgt, DIMS)[0]) / (2 * h)
    assert abs(fd - g[i]) < 1e-6 * max(1, abs(fd))
    assert abs(fd + g[i]) > 1e-3 * abs(fd)


def test_spectral_identity_on_band_limited_fields():
    """With R = 1 and the Hermitian weights, E(s D z) = z for every real z
    whose modes satisfy |kx| < M and |ky| < M; modes outside are removed."""
    M = 5
    y, x = np.meshgrid(np.arange(N) * 2 * np.pi / N, np.arange(N) * 2 * np.pi / N, indexing="ij")
    z = 0.3 + np.cos(2 * x + 3 * y) + np.sin(x - 4 * y) - 0.7 * np.cos(4 * x) + 0.2 * np.sin(3 * y)
    s = F.mode_scale(M, N)
    assert np.allclose(F.idft_re(s * F.dft(z, M), N), z, atol=1e-12)
    hi = np.cos(6 * x + y) + np.sin(x + 7 * y)
    assert np.allclose(F.idft_re(s * F.dft(z + hi, M), N), z, atol=1e-12)


def test_dft_matches_brute_force():
    M = 3
    rng = np.random.default_rng(4)
    z = rng.normal(0, 1, (N, N))
    X = F.dft(z, M)
    ys, xs = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    for kx in range(M):
        for j in range(2 * M):
            ky = j if j < M else j - 2 * M
            ref = np.sum(z * np.exp(-2j * np.pi * (kx * xs + ky * ys) / N))
            assert abs(X[kx, j] - ref) < 1e-10


def test_bf16_rounding():
    a = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 1 + 2 ** -7 + 2 ** -9], dtype=np.float32)
    r = F.bf16_round(a)
    assert r.tolist() == [1.0, 1.0, 1 + 2 ** -6, -2.5, 1 + 2 ** -7]


def test_adam_first_step_is_signed_lr():
    w = np.array([0.5, -0.25, 2.0])
    g = np.array([0.3, -4.0, 1e-3])
    w1, m, v = F.adam(w, np.zeros(3), np.zeros(3), g, 1, 1e-2)
    assert np.allclose(w1, w - 1e-2 * np.sign(g), rtol=0, atol=1e-7)


def test_layout_counts():
    cin, C, M, L, P = 2, 32, 16, 4, 64
    _, n = F.layout(cin, C, M, L, P)
    assert n == C * cin + C + L * (2 * C * C * 2 * M * M + C * C + C) + P * C + P + 2 * P + 2
