! A reference-style particle run on the ISO_C_BINDING layer, in the order of
! src/driver.f90:460-508 and src/solve.f90:105-136: namelist globals, the
! argument-less initialize_fft, restart_gpe + the ARGLE-data sync, then
! particle_potential and initialize_particles from restart/part files, RK4
! steps with diagnostics_gpe every 5 steps (spectra every 10), and
! gppart_checkpoint at the end.  Argument 1: run directory holding
! restart/wf/wf0p1.dat and restart/part/gp*.dat_nrst7; data/ and spectra/
! are written there, and the final psi, kpsi and particle state to
! final.bin.
PROGRAM driver_part
  USE ISO_C_BINDING
  USE vikshobha_c
  USE vikshobha_gpu
  IMPLICIT NONE
  INTEGER, PARAMETER :: nsteps = 20
  INTEGER :: itime, u
  REAL(C_DOUBLE) :: pi, time
  CHARACTER(256) :: rundir

  CALL GET_COMMAND_ARGUMENT(1, rundir)
  pi = 4.0_C_DOUBLE * ATAN(1.0_C_DOUBLE)
  ! namelist inputs (src/driver.f90:52-78)
  Nx_in = 64
  dt_fix = 1.0e-3_C_DOUBLE
  xi = 2.5_C_DOUBLE
  csound = SQRT(2.0_C_DOUBLE) * xi * (2.0_C_DOUBLE * pi / Nx_in)
  rho_avg = 1.0_C_DOUBLE
  disp_factor = 0.0_C_DOUBLE
  gp_N_obj = 3
  ppotV0 = 10.0_C_DOUBLE
  ppotd = 2.0_C_DOUBLE
  gp_mass_part = 5.0_C_DOUBLE
  nrst = 7
  gsteps = 5
  ssteps = 10
  wf_in_dir = TRIM(rundir)//'/restart/wf'
  part_in_dir = TRIM(rundir)//'/restart/part'
  data_dir = TRIM(rundir)//'/data'
  spectra_dir = TRIM(rundir)//'/spectra'

  CALL initialize_fft()
  CALL restart_gpe(0)
  CALL set_initial_psi()
  CALL particle_potential()
  CALL initialize_particles()
  time = 0.0_C_DOUBLE
  DO itime = 1, nsteps
    time = time + dt_fix
    CALL gpe_rk4()
    IF (MOD(itime, 5) == 0) CALL diagnostics_gpe(itime / ssteps, itime, time)
  END DO
  CALL gppart_checkpoint(1, nsteps, time)
  CALL sync_host()
  OPEN(NEWUNIT=u, FILE=TRIM(rundir)//'/final.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) psi, kpsi, gp_pos_obj, gp_vel_obj, gp_frc_obj, eppot, eobj_kin, esrep, mom_obj
  CLOSE(u)
  WRITE(*,'(A,ES22.14,A,ES22.14)') 'driver_part: norm=', norm, ' eppot=', eppot
  CALL clear_fft()
END PROGRAM driver_part
