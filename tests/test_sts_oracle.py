"""Spatio-temporal spectra: the numpy oracle pinned by mathematics, and the
host-side file formats of sts_gpe (src/diagnostics.f90:636-695) checked
against a Fortran runtime reading them with CompSTS's own READ statements
(Run/CompSTS/driver.f90:141-151, 188-200)."""
import os
import subprocess

import numpy as np
import pytest

import sts_oracle as S

FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def _snaps(nt, N, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nt, N, N)) + 1j * rng.standard_normal((nt, N, N))


@pytest.mark.parametrize("nt", [32, 64, 256, 4096])
def test_hann_closed_form(nt):
    w = S.hann(nt)
    # sum_{n=0}^{Nt-1} cos(2 pi n/(Nt-1)) = 1 (a full period plus the endpoint)
    assert abs(w.sum() - (nt - 1) / 2.0) < 1e-12 * nt
    assert w[0] == 0.0 and abs(w[-1]) < 1e-15
    assert np.allclose(w, w[::-1], atol=1e-15, rtol=0)


def test_time_transform_is_the_forward_dft():
    x = _snaps(32, 4, 1)
    F = np.fft.fft(S.hann(32)[:, None, None] * x, axis=0)
    B = S.dft_bruteforce(S.hann(32)[:, None, None] * x)
    assert np.max(np.abs(F - B)) < 1e-12 * np.max(np.abs(B))
    sp2, sp1 = S.sts_spectra(x)
    P = (1 / 32) ** 2 * (1 / 16) ** 2 * np.abs(B) ** 2
    assert np.allclose(sp1, P[:, 0, :], rtol=1e-12, atol=0)
    assert np.allclose(sp2.sum(axis=1), P.sum(axis=(1, 2)), rtol=1e-12, atol=0)


@pytest.mark.parametrize("kx,ky,j", [(3, 4, 5), (-3, 4, 17), (0, -7, 60), (5, 12, 0)])
def test_single_tone(kx, ky, j):
    """x_n = e^{+2 pi i j n/Nt} on one mode: the forward DFT at w = j is the
    window sum (Nt-1)/2, so spec2d[j, nint|k|] = stsnorm ((Nt-1)/2)^2 and
    every other shell is empty."""
    nt, N = 64, 32
    x = np.zeros((nt, N, N), dtype=np.complex128)
    x[:, kx % N, ky % N] = np.exp(2j * np.pi * j * np.arange(nt) / nt)
    sp2, sp1 = S.sts_spectra(x)
    sh = int(round(np.hypot(kx, ky)))
    stsnorm = (1 / nt) ** 2 * (1 / N ** 2) ** 2
    assert abs(sp2[j, sh] - stsnorm * ((nt - 1) / 2) ** 2) < 1e-12 * sp2[j, sh]
    assert np.argmax(sp2[:, sh]) == j
    others = np.delete(sp2, sh, axis=1)
    assert np.all(others == 0.0)
    if kx == 0:
        assert np.array_equal(sp1[:, ky % N], sp2[:, sh])
    else:
        assert np.all(sp1 == 0.0)


def test_parseval():
    nt, N = 128, 16
    x = _snaps(nt, N, 2)
    sp2, _ = S.sts_spectra(x)
    stsnorm = (1 / nt) ** 2 * (1 / N ** 2) ** 2
    w = S.hann(nt)
    lhs = sp2.sum()
    rhs = stsnorm * nt * np.sum(w[:, None, None] ** 2 * np.abs(x) ** 2)
    assert abs(lhs - rhs) < 1e-12 * rhs


def test_shell_index():
    sh = S.shell_index(16)
    assert sh[0, 0] == 0 and sh[3, 4] == 5 and sh[16 - 3, 4] == 5 and sh[8, 8] == 11
    # Nyquist index maps to -N/2
    assert sh[8, 0] == 8
    assert sh.max() <= S.nshell(16)


def test_fortran_e_format():
    from quantumvortices_amd.gpe import _fortran_e
    assert _fortran_e(1.0) == "   0.100000E+01"
    assert _fortran_e(123.456) == "   0.123456E+03"
    assert _fortran_e(9.9999999) == "   0.100000E+02"
    assert _fortran_e(-2.5e-7) == "  -0.250000E-06"
    assert _fortran_e(0.0) == "   0.000000E+00"


FORTRAN_STS = """
program stscheck
  implicit none
  integer, parameter :: Nx = 16, Ny = 16, Nt = 4, nsplit = 2, L = Nx / nsplit
  integer :: i1, i2, i3, i4, ios
  complex(kind=8) :: x0(Ny, Nt), pl(Ny, Nx, Nt)
  double precision :: s0, s2, t
  character(len=8) :: prcjj
  ! Run/CompSTS/driver.f90: READ(1) (sts1d1(i2,i4),i2=1,Ny) per snapshot
  open(unit=1, file='sts1d/kpsi_x0_nrst3.dat', form='UNFORMATTED', status='OLD', iostat=ios)
  do i4 = 1, Nt
    read(1) (x0(i2, i4), i2 = 1, Ny)
  end do
  close(1)
  ! and the slabs of planes, kx in ((i3-1)*L+1 .. i3*L), ky fastest
  do i3 = 1, nsplit
    write(prcjj, '(i8)') i3
    open(unit=1, file='sts2d/kpsi_nrst3p'//trim(adjustl(prcjj))//'.dat', form='UNFORMATTED', status='OLD', iostat=ios)
    do i4 = 1, Nt
      read(1) ((pl(i2, i1, i4), i2 = 1, Ny), i1 = (i3 - 1) * L + 1, i3 * L)
    end do
    close(1)
  end do
  s0 = 0.0d0
  s2 = 0.0d0
  do i4 = 1, Nt
    do i2 = 1, Ny
      s0 = s0 + dble(x0(i2, i4)) * i2 + aimag(x0(i2, i4)) * i4
      do i1 = 1, Nx
        s2 = s2 + dble(pl(i2, i1, i4)) * (i1 + 100 * i2) + aimag(pl(i2, i1, i4)) * (i4 + 7 * i1)
      end do
    end do
  end do
  open(unit=1, file='sts_time.dat_nrst3', status='OLD')
  do i4 = 1, Nt
    read(1, '(E15.6)') t
    s0 = s0 + t
  end do
  close(1)
  print '(2ES25.16)', s0, s2
end program
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_sts_files_read_by_fortran(tmp_path):
    from quantumvortices_amd import gpe as G
    N, nt = 16, 4
    rng = np.random.default_rng(3)
    planes = rng.standard_normal((nt, N, N)) + 1j * rng.standard_normal((nt, N, N))
    x0 = planes[:, 0, :]
    times = [0.25, 0.5, 0.75, 1.0]
    G.write_sts_files(tmp_path, 3, times, x0, planes, N, nsplit=2)
    (tmp_path / "stscheck.f90").write_text(FORTRAN_STS)
    r = subprocess.run([FC, "-O0", "-o", str(tmp_path / "stscheck"), str(tmp_path / "stscheck.f90")],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(tmp_path / "stscheck")], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stderr
    s0, s2 = (float(v) for v in r.stdout.split())
    i2 = np.arange(1, N + 1)
    i4 = np.arange(1, nt + 1)
    e0 = np.sum(x0.real * i2[None, :] + x0.imag * i4[:, None]) + sum(times)
    # planes[n][kx][ky] = pl(ky+1, kx+1, n+1)
    i1 = np.arange(1, N + 1)
    e2 = np.sum(planes.real * (i1[None, :, None] + 100 * i2[None, None, :]) +
                planes.imag * (i4[:, None, None] + 7 * i1[None, :, None]))
    assert abs(s0 - e0) < 1e-12 * abs(e0) and abs(s2 - e2) < 1e-12 * abs(e2)
