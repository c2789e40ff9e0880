"""Checkpoint files in the reference's formats (host code, no GPU):
wf files of io_fields / restart_gpe (src/io_fields.f90) and the particle
text files (src/gpe_particle.f90:92-184).  The binary layout is pinned
against a Fortran runtime: amdflang writes a wf file with the reference's
own WRITE statements and it must equal ours byte for byte, and a Fortran
READ in the reference's order must read ours back."""
import os
import shutil
import subprocess

import numpy as np
import pytest

FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def _psi(N, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))


def test_wf_round_trip_and_layout(tmp_path):
    from quantumvortices_amd import gpe as G
    N = 32
    psi = _psi(N)
    G.write_wf(tmp_path, 7, 1.25, 2.5e-4, psi)
    icp, t, dt, back = G.read_wf(tmp_path, 7, N)
    assert (icp, t, dt) == (7, 1.25, 2.5e-4)
    assert np.array_equal(back, psi)
    raw = (tmp_path / "wf7p1.dat").read_bytes()
    # header record: 20-byte payload framed by its length
    assert len(raw) == (4 + 20 + 4) + N * (4 + 16 * N + 4)
    assert np.frombuffer(raw[:4], "<i4")[0] == 20 and np.frombuffer(raw[24:28], "<i4")[0] == 20
    assert np.frombuffer(raw[4:8], "<i4")[0] == 7
    assert np.frombuffer(raw[8:24], "<f8").tolist() == [1.25, 2.5e-4]
    # first row: x fastest, (re, im) pairs
    row = np.frombuffer(raw[32:32 + 16 * N], "<f8").reshape(N, 2)
    assert np.array_equal(row[:, 0], psi[0].real) and np.array_equal(row[:, 1], psi[0].imag)


@pytest.mark.parametrize("nsplit", [2, 4, 8])
def test_wf_regrouping(tmp_path, nsplit):
    """Files written by nsplit slabs (ranks) regroup into one field; each file
    carries N/nsplit rows in rank order."""
    from quantumvortices_amd import gpe as G
    N = 64
    psi = _psi(N, 1)
    G.write_wf(tmp_path, 3, 0.5, 1e-3, psi, nsplit=nsplit)
    files = sorted(os.listdir(tmp_path))
    assert files == sorted("wf3p%d.dat" % (r + 1) for r in range(nsplit))
    for r in range(nsplit):
        raw = (tmp_path / ("wf3p%d.dat" % (r + 1))).read_bytes()
        assert len(raw) == 28 + (N // nsplit) * (8 + 16 * N)
    _, _, _, back = G.read_wf(tmp_path, 3, N, nsplit_orig=nsplit)
    assert np.array_equal(back, psi)
    with pytest.raises(Exception):
        G.read_wf(tmp_path, 3, N, nsplit_orig=nsplit * 2)   # missing files / wrong record sizes


FORTRAN_WF = """
program wfcheck
  implicit none
  integer, parameter :: Nx = 16, Ny = 16, nsplit = 2
  integer :: i1, i2, i2_loc, r, iopoint, icp
  double precision :: time, dt, tr, dtr, s
  complex(kind=8) :: psi(Nx, Ny), back(Nx, Ny)
  character(len=8) :: fnum, prc
  do i2 = 1, Ny
    do i1 = 1, Nx
      psi(i1, i2) = cmplx(i1 + 1000.0d0 * i2, -0.5d0 * i1 + 1.0d0 / i2, kind=8)
    end do
  end do
  iopoint = 5; time = 0.375d0; dt = 1.0d-3
  ! src/io_fields.f90:41-50, one file per slab
  write(fnum, '(i8)') iopoint
  do r = 1, nsplit
    write(prc, '(i8)') r
    open(unit=18, file='fort/wf'//trim(adjustl(fnum))//'p'//trim(adjustl(prc))//'.dat', &
         form='UNFORMATTED', status='NEW')
    write(18) iopoint, time, dt
    do i2_loc = 1, Ny / nsplit
      write(18) (psi(i1, (r - 1) * (Ny / nsplit) + i2_loc), i1 = 1, Nx)
    end do
    close(18)
  end do
  ! src/io_fields.f90:100-120 (regrouping read) on the files the library wrote
  do r = 1, nsplit
    write(prc, '(i8)') r
    open(unit=11, file='ours/wf'//trim(adjustl(fnum))//'p'//trim(adjustl(prc))//'.dat', &
         form='UNFORMATTED', status='OLD')
    read(11) icp, tr, dtr
    do i2_loc = 1, Ny / nsplit
      read(11) (back(i1, (r - 1) * (Ny / nsplit) + i2_loc), i1 = 1, Nx)
    end do
    close(11)
  end do
  s = maxval(abs(back - psi))
  print '(I6,2ES25.16,ES12.3)', icp, tr, dtr, s
end program
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_wf_matches_fortran_runtime(tmp_path):
    from quantumvortices_amd import gpe as G
    Nx = 16
    i1 = np.arange(1, Nx + 1)
    psi = np.array([[complex(a + 1000.0 * b, -0.5 * a + 1.0 / b) for a in i1] for b in i1])   # [y, x]
    (tmp_path / "fort").mkdir()
    (tmp_path / "ours").mkdir()
    G.write_wf(tmp_path / "ours", 5, 0.375, 1e-3, psi, nsplit=2)
    (tmp_path / "wfcheck.f90").write_text(FORTRAN_WF)
    r = subprocess.run([FC, "-O0", "-o", str(tmp_path / "wfcheck"), str(tmp_path / "wfcheck.f90")],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(tmp_path / "wfcheck")], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stderr
    icp, tr, dtr, err = r.stdout.split()
    assert int(icp) == 5 and float(tr) == 0.375 and float(dtr) == 1e-3 and float(err) == 0.0
    for p in (1, 2):
        assert (tmp_path / "fort" / ("wf5p%d.dat" % p)).read_bytes() == (tmp_path / "ours" / ("wf5p%d.dat" % p)).read_bytes()
    # and the library reads what Fortran wrote
    icp, t, dt, back = G.read_wf(tmp_path / "fort", 5, Nx, nsplit_orig=2)
    assert (icp, t, dt) == (5, 0.375, 1e-3) and np.array_equal(back, psi)


FORTRAN_PART = """
program partcheck
  implicit none
  integer :: i
  double precision :: p(2, 3)
  p = reshape([1.0d0/3.0d0, 2.5d0, -7.25d-12, 6.283185307179586d0, 1.0d300, 0.1d0], [2, 3])
  open(unit=1, file='gppos.dat_nrst4')
  do i = 1, 3
    write(1, *) 12, 0.5d0, p(1, i), p(2, i)
  end do
  close(1)
end program
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_particle_files_read_fortran_list_directed(tmp_path):
    """A position file written by Fortran list-directed WRITE(1,*) (as
    gppart_checkpoint does) reads back exactly."""
    from quantumvortices_amd import _native as nat
    (tmp_path / "partcheck.f90").write_text(FORTRAN_PART)
    r = subprocess.run([FC, "-O0", "-o", str(tmp_path / "partcheck"), str(tmp_path / "partcheck.f90")],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([str(tmp_path / "partcheck")], cwd=tmp_path, timeout=60).returncode == 0
    for name in ("gpvel", "gpf0ext"):
        shutil.copy(tmp_path / "gppos.dat_nrst4", tmp_path / ("%s.dat_nrst4" % name))
    pos, vel, fext = np.zeros((3, 2)), np.zeros((3, 2)), np.zeros((3, 2))
    nat.check(nat.lib().vk_part_restart(str(tmp_path).encode(), 4, 3, nat.ptr(pos), nat.ptr(vel), nat.ptr(fext)))
    ref = np.array([[1.0 / 3.0, 2.5], [-7.25e-12, 6.283185307179586], [1e300, 0.1]])
    assert np.array_equal(pos, ref) and np.array_equal(fext, ref)


def test_particle_checkpoint_round_trip(tmp_path):
    from quantumvortices_amd import _native as nat
    rng = np.random.default_rng(2)
    pos, vel, fext = (rng.standard_normal((5, 2)) for _ in range(3))
    for k in range(2):   # appends: the restart reads the first np lines
        nat.check(nat.lib().vk_part_checkpoint(str(tmp_path).encode(), 0, 10 + k, 0.25 * k, 5,
                                               nat.ptr(pos), nat.ptr(vel), nat.ptr(fext)))
    lines = (tmp_path / "gppos_ckp.dat_nrst0").read_text().splitlines()
    assert len(lines) == 10 and lines[0].split()[0] == "10" and lines[5].split()[0] == "11"
    for a in ("gppos", "gpvel", "gpf0ext"):
        shutil.copy(tmp_path / ("%s_ckp.dat_nrst0" % a), tmp_path / ("%s.dat_nrst0" % a))
    p2, v2, f2 = np.zeros((5, 2)), np.zeros((5, 2)), np.zeros((5, 2))
    nat.check(nat.lib().vk_part_restart(str(tmp_path).encode(), 0, 5, nat.ptr(p2), nat.ptr(v2), nat.ptr(f2)))
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel) and np.array_equal(f2, fext)
