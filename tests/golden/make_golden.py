"""Writes the FFT golden fixtures (numpy/pocketfft) used to pin the oracle's
and the library's 2D transform conventions.  The reference's transforms are
FFTW3-MPI (not vendored, not importable here): its convention -- sign -1
forward, unnormalised, k-space stored transposed [kx][ky], inverse scaled by
1/(Nx*Ny) by the caller (src/mod_fft.f90:307-311,414,482-487) -- is restated
on numpy.fft.fft2 of the x-space array [y][x]."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for N in (8, 16, 32):
        rng = np.random.default_rng(1000 + N)
        a = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        fwd = np.fft.fft2(a).T                       # k-space [kx, ky]
        inv = np.fft.ifft2(a.T)                      # input read as k-space [kx,ky] -> x-space [y,x]
        np.savez(os.path.join(HERE, "fft_N%d.npz" % N), x=a, fwd=fwd, inv_in=a, inv=inv,
                 numpy_version=np.__version__)


if __name__ == "__main__":
    main()
