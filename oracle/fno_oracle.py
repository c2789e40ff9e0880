"""CPU oracle of the Fourier neural operator used for the ARGLE warm start.

TEST INFRASTRUCTURE ONLY: imported by tests/ and nothing in the product path.
Float64 numpy restatement of the network, its loss gradient and one Adam
step.  There is no reference counterpart (the reference starts ARGLE from
the analytic ansatz of src/initial_config.f90), so parity is pinned by
mathematics instead of reference vectors:
  * the gradient against central finite differences (tests/test_fno_oracle.py);
  * the spectral convolution with R = 1 against the identity on band-limited
    fields, which fixes the mode set, the Hermitian weights and the 1/N^2;
  * D against numpy's FFT (it is computed with np.fft.fft2 here, while the
    GPU evaluates separable sums over e^{-2 pi i k x / N} tables).

Parameter vector layout (floats), identical to vk_fno_layout:
  Wp[C][cin], bp[C], per layer l: R_l[C_in][C_out][M][2M] complex as
  (re, im) pairs, W_l[C_out][C_in], b_l[C], then Q1[P][C], bq1[P], Q2[2][P],
  bq2[2].  Mode (kx, j): kx in [0, M), ky = j for j < M else j - 2M.
"""
import numpy as np
from scipy.special import erf


def layout(cin, C, M, L, P):
    """[(name, offset, shape)] and the total count."""
    out = []
    o = 0

    def add(name, shape, n):
        nonlocal o
        out.append((name, o, shape))
        o += n

    add("Wp", (C, cin), C * cin)
    add("bp", (C,), C)
    for l in range(L):
        add("R%d" % l, (C, C, M, 2 * M, 2), C * C * 2 * M * M * 2)
        add("W%d" % l, (C, C), C * C)
        add("b%d" % l, (C,), C)
    add("Q1", (P, C), P * C)
    add("bq1", (P,), P)
    add("Q2", (2, P), 2 * P)
    add("bq2", (2,), 2)
    return out, o


def unpack(p, cin, C, M, L, P):
    lay, n = layout(cin, C, M, L, P)
    assert p.size == n
    w = {}
    for name, off, shape in lay:
        size = int(np.prod(shape))
        a = np.asarray(p[off:off + size], dtype=np.float64).reshape(shape)
        if name.startswith("R"):
            a = a[..., 0] + 1j * a[..., 1]
        w[name] = a
    return w


def pack_grad(g, cin, C, M, L, P):
    lay, n = layout(cin, C, M, L, P)
    out = np.zeros(n)
    for name, off, shape in lay:
        a = g[name]
        if name.startswith("R"):
            a = np.stack([a.real, a.imag], axis=-1)
        out[off:off + a.size] = a.ravel()
    return out


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def _jidx(M, N):
    return np.array([j if j < M else N + j - 2 * M for j in range(2 * M)])


def dft(z, M):
    """D z: (..., N, N)[y][x] -> (..., M, 2M)[kx][j], unnormalised forward."""
    N = z.shape[-1]
    Z = np.fft.fft2(z, axes=(-2, -1))           # Z[..., ky, kx]
    return np.swapaxes(Z[..., _jidx(M, N), :M], -1, -2)


def idft_re(Y, N):
    """E Y = Re sum_{kx, ky} Y e^{+2 pi i (kx x + ky y)/N} -> (..., N, N)."""
    M = Y.shape[-2]
    full = np.zeros(Y.shape[:-2] + (N, N), dtype=np.complex128)
    full[..., _jidx(M, N), :M] = np.swapaxes(Y, -1, -2)
    return (np.fft.ifft2(full, axes=(-2, -1)) * (N * N)).real


def mode_scale(M, N):
    s = np.full((M, 1), 2.0 / (N * N))
    s[0, 0] = 1.0 / (N * N)
    return s                                     # broadcasts over [kx][j]


def forward(p, x, dims):
    """x: (B, cin, N, N) -> out (B, 2, N, N), cache."""
    cin, C, M, L, P = dims
    w = unpack(p, cin, C, M, L, P)
    N = x.shape[-1]
    s = mode_scale(M, N)
    z = [np.einsum("ck,bkyx->bcyx", w["Wp"], x) + w["bp"][None, :, None, None]]
    us, Xs = [], []
    for l in range(L):
        X = dft(z[l], M)                                            # (B, C, M, 2M)
        Ym = s * np.einsum("bimj,iomj->bomj", X, w["R%d" % l])
        u = (np.einsum("oi,biyx->boyx", w["W%d" % l], z[l]) + w["b%d" % l][None, :, None, None]
             + idft_re(Ym, N))
        Xs.append(X)
        us.append(u)
        z.append(gelu(u) if l < L - 1 else u)
    h = np.einsum("pc,bcyx->bpyx", w["Q1"], z[L]) + w["bq1"][None, :, None, None]
    out = np.einsum("qp,bpyx->bqyx", w["Q2"], gelu(h)) + w["bq2"][None, :, None, None]
    return out, dict(w=w, z=z, u=us, X=Xs, h=h, x=x)


def loss_grad(p, x, tgt, dims):
    """Mean squared error and its gradient (flat, same layout as p)."""
    cin, C, M, L, P = dims
    out, c = forward(p, x, dims)
    w, z, us, Xs, h = c["w"], c["z"], c["u"], c["X"], c["h"]
    N = x.shape[-1]
    s = mode_scale(M, N)
    n = out.size
    d = out - tgt
    loss = float(np.sum(d * d) / n)
    dout = 2.0 * d / n
    g = {}
    gh = gelu(h)
    g["Q2"] = np.einsum("bqyx,bpyx->qp", dout, gh)
    g["bq2"] = dout.sum(axis=(0, 2, 3))
    dh = np.einsum("qp,bqyx->bpyx", w["Q2"], dout) * gelu_d(h)
    g["Q1"] = np.einsum("bpyx,bcyx->pc", dh, z[L])
    g["bq1"] = dh.sum(axis=(0, 2, 3))
    dz = np.einsum("pc,bpyx->bcyx", w["Q1"], dh)
    for l in range(L - 1, -1, -1):
        du = dz * gelu_d(us[l]) if l < L - 1 else dz
        g["W%d" % l] = np.einsum("boyx,biyx->oi", du, z[l])
        g["b%d" % l] = du.sum(axis=(0, 2, 3))
        Yb = dft(du, M)                                             # (B, C_out, M, 2M)
        g["R%d" % l] = s * np.einsum("bomj,bimj->iomj", Yb, np.conj(Xs[l]))
        Zi = s * np.einsum("bomj,iomj->bimj", Yb, np.conj(w["R%d" % l]))
        dz = np.einsum("oi,boyx->biyx", w["W%d" % l], du) + idft_re(Zi, N)
    g["Wp"] = np.einsum("bcyx,bkyx->ck", dz, x)
    g["bp"] = dz.sum(axis=(0, 2, 3))
    return loss, pack_grad(g, cin, C, M, L, P)


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), as floats."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)
