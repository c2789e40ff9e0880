! A reference-style driver on the ISO_C_BINDING layer: sets up a 4x4 Padé
! vortex lattice at N=64, syncs it like the reference's restart path, runs
! RK4 steps and diagnostics through the reference's subroutine names, and
! writes psi, kpsi and the energies to a stream file (argument 1) for
! comparison, the diagnostics' text files into argument 2; then
! takes Ntt = 32 spatio-temporal snapshots, one per RK4 step, and appends
! spec_sts2d and arr1; then transforms psi with the work-array plans of
! fft() ('gpnlin' forward and back, 'cmpxtmp' forward on psi**2) and
! appends psi, the transforms and the round trip.
PROGRAM driver_check
  USE ISO_C_BINDING
  USE vikshobha_c
  USE vikshobha_gpu
  IMPLICIT NONE
  INTEGER, PARAMETER :: N = 64, nsteps = 10
  REAL(C_DOUBLE) :: pi, dx, xi_l, cs_l, xs(16), ys(16)
  INTEGER(C_INT) :: qs(16), rc
  INTEGER :: i, j, k, u
  CHARACTER(256) :: outname, outdir

  pi = 4.0_C_DOUBLE * ATAN(1.0_C_DOUBLE)
  dx = 2.0_C_DOUBLE * pi / N
  xi_l = 2.5_C_DOUBLE
  cs_l = SQRT(2.0_C_DOUBLE) * xi_l * dx
  CALL GET_COMMAND_ARGUMENT(1, outname)
  CALL GET_COMMAND_ARGUMENT(2, outdir)
  data_dir = TRIM(outdir)
  spectra_dir = TRIM(outdir)
  CALL initialize_fft(N, 1.0e-3_C_DOUBLE, xi_l, cs_l, 1.0_C_DOUBLE, 0.0_C_DOUBLE)
  k = 0
  DO i = 0, 3
    DO j = 0, 3
      k = k + 1
      xs(k) = (i + 0.5_C_DOUBLE) * 2.0_C_DOUBLE * pi / 4 + 0.31_C_DOUBLE * dx
      ys(k) = (j + 0.5_C_DOUBLE) * 2.0_C_DOUBLE * pi / 4 + 0.17_C_DOUBLE * dx
      qs(k) = MERGE(1, -1, MOD(i + j, 2) == 0)
    END DO
  END DO
  rc = vk_ic_vortices_host(INT(N, C_INT), 16_C_INT, xs, ys, qs, params%xi, 3_C_INT, psi)
  CALL vk_check(rc, 'ic')
  CALL set_initial_psi()
  DO i = 1, nsteps
    CALL gpe_rk4()
  END DO
  CALL diagnostics_gpe(0, nsteps, nsteps * 1.0e-3_C_DOUBLE)
  CALL sync_host()
  OPEN(NEWUNIT=u, FILE=TRIM(outname), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) psi, kpsi, egrad, echem, eintr, etot, norm, momx, momy, spec_gp
  Ntt = 32
  DO i = 1, Ntt
    CALL gpe_rk4()
    CALL sts_gpe(i, (nsteps + i) * 1.0e-3_C_DOUBLE)
  END DO
  CALL comp_sts()
  WRITE(u) spec_sts2d, arr1
  ! the work-array plans of fft(): gpnlin forward and back, cmpxtmp forward
  CALL sync_host()
  cgp_phys_nlin = psi
  CALL fft('gpnlin', forward)
  WRITE(u) psi, cgp_four_nlin
  CALL fft('gpnlin', inverse)
  WRITE(u) cgp_phys_nlin
  phys_cmpxtmp = psi * psi
  CALL fft('cmpxtmp', forward)
  WRITE(u) four_cmpxtmp
  CLOSE(u)
  WRITE(*,'(A,ES22.14,A,ES22.14)') 'driver_check: norm=', norm, ' etot=', etot
  CALL clear_fft()
END PROGRAM driver_check
