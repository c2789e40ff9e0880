"""CPU oracle of the spatio-temporal spectra (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and nothing in the product path.
Restates the post-processing of Run/CompSTS/driver.f90:134-265 on the
snapshots sts_gpe writes (src/diagnostics.f90:636-695), with numpy's FFT in
place of FFTW's forward DFT (same sign e^{-2 pi i n w / Nt}, no scaling):

  w_n      = (1 - cos(2 pi n / (Nt - 1))) / 2,  n = 0 .. Nt-1   (:163-165, :219-221)
  stsnorm  = (1/Nt)^2 (1/(Nx Ny))^2                            (driver.f90 header)
  spec1d[w][ky]  = stsnorm |DFT_n(w_n kpsi_n(ky, kx=0))[w]|^2      (:159-173)
  spec2d[w][sh]  = sum over all (kx, ky) with nint(|k|) = sh of
                   stsnorm |DFT_n(w_n kpsi_n(ky, kx))[w]|^2        (:229-258)

Snapshots are kpsi_t[kx, ky] planes (kx slow, ky fast: the reference's
kpsi(ky, kx) in column-major order).  Pinned by tests/test_sts_oracle.py
against mathematics: a brute-force DFT, the closed form of the window sum
((Nt - 1)/2), a single tone on one mode (all power in one shell and one
frequency, of value stsnorm ((Nt-1)/2)^2), and Parseval.
"""
import numpy as np


def hann(nt):
    n = np.arange(nt, dtype=np.float64)
    return 0.5 * (1.0 - np.cos(2.0 * np.pi * n / (nt - 1)))


def nshell(N):
    """Upper shell index of spec_sts2d(0:nshell, :) (driver.f90 header)."""
    return int(1.4142 * N / 2.0) + 1


def shell_index(N):
    """mshl = NINT(SQRT(k1^2 + k2^2)) per (kx, ky) index; k = (i-1) - N*(i/(N/2+1))."""
    i = np.arange(N)
    k = np.where(i < N // 2, i, i - N)
    kk = k[:, None] ** 2 + k[None, :] ** 2
    return np.rint(np.sqrt(kk.astype(np.float64))).astype(np.int64)


def sts_spectra(snaps):
    """snaps: (Nt, N, N) complex, kpsi_t[kx, ky] per snapshot.
    Returns spec2d (Nt, nshell(N)+1) and spec1d (Nt, N)."""
    snaps = np.asarray(snaps, dtype=np.complex128)
    nt, N, _ = snaps.shape
    stsnorm = (1.0 / nt) ** 2 * (1.0 / (N * N)) ** 2
    F = np.fft.fft(hann(nt)[:, None, None] * snaps, axis=0)
    P = stsnorm * (F.real ** 2 + F.imag ** 2)
    spec1d = P[:, 0, :].copy()
    sh = shell_index(N).ravel()
    spec2d = np.zeros((nt, nshell(N) + 1))
    for w in range(nt):
        spec2d[w] = np.bincount(sh, weights=P[w].ravel(), minlength=nshell(N) + 1)
    return spec2d, spec1d


def dft_bruteforce(x):
    """Forward DFT along axis 0 by the defining sum (small Nt only)."""
    nt = x.shape[0]
    n = np.arange(nt)
    E = np.exp(-2j * np.pi * np.outer(n, n) / nt)
    return np.tensordot(E, x, axes=(1, 0))
