! ISO_C_BINDING layer over the vikshobha C ABI (include/vikshobha.h), keeping
! the reference's module API so a Fortran driver can call it in place of
! mod_fft / gpe / diagnostics:
!   initialize_fft            <- src/mod_fft.f90:288 (plans) + create_state_variables :122
!                                (argument-less on the namelist globals, as in
!                                the reference, or with explicit arguments)
!   fft(plan_name, direction) <- src/mod_fft.f90:471  (plan 'psi')
!   gpe_rk4                   <- src/gpe.f90:310
!   ARGLE(time)               <- src/gpe.f90:5
!   diagnostics_gpe(spoint,jtime,time) <- src/diagnostics.f90:4, with the
!                                files of :240-259, :530-569 (particles),
!                                quantglobal_gpe :573 and spectrum_gpe :608
!   particle_potential        <- src/gpe_particle.f90:3
!   initialize_particles      <- src/gpe_particle.f90:92 (restart/part files)
!   gppart_checkpoint(iopoint,jtime,ptime) <- src/gpe_particle.f90:143
!   io_fields(iopoint,time)   <- src/io_fields.f90:1 (wf files, IOEVERYPROC layout)
!   restart_gpe(nrestart)     <- src/io_fields.f90:71 (with nsplit regrouping)
!   clear_fft                 <- src/mod_fft.f90:433
! psi(Nx,Ny) and kpsi(Ny,Nx) are host arrays with the reference's layouts.
! The state lives on the device between calls; psi/kpsi are refreshed by
! fft(...) and by sync_host().

MODULE vikshobha_c
  USE ISO_C_BINDING
  IMPLICIT NONE

  TYPE, BIND(C) :: vk_gpe_params
    INTEGER(C_INT) :: N
    REAL(C_DOUBLE) :: dt, alpha, g, omega, disp, xi
    INTEGER(C_INT) :: np
    REAL(C_DOUBLE) :: V0, d, mass
    INTEGER(C_INT) :: force_dealias
  END TYPE

  INTERFACE
    INTEGER(C_INT) FUNCTION vk_gpe_params_init(p, N, dt, xi_in, csound, rho_avg, disp) BIND(C, NAME='vk_gpe_params_init')
      IMPORT :: C_INT, C_DOUBLE, vk_gpe_params
      TYPE(vk_gpe_params) :: p
      INTEGER(C_INT), VALUE :: N
      REAL(C_DOUBLE), VALUE :: dt, xi_in, csound, rho_avg, disp
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_nshell(N) BIND(C, NAME='vk_nshell')
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: N
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_create(h, p) BIND(C, NAME='vk_gpe_create')
      IMPORT :: C_INT, C_PTR, vk_gpe_params
      TYPE(C_PTR) :: h
      TYPE(vk_gpe_params) :: p
    END FUNCTION
    SUBROUTINE vk_gpe_destroy(h) BIND(C, NAME='vk_gpe_destroy')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
    END SUBROUTINE
    INTEGER(C_INT) FUNCTION vk_gpe_set_psi(h, psi, flags) BIND(C, NAME='vk_gpe_set_psi')
      IMPORT :: C_INT, C_PTR, C_DOUBLE_COMPLEX
      TYPE(C_PTR), VALUE :: h
      COMPLEX(C_DOUBLE_COMPLEX) :: psi(*)
      INTEGER(C_INT), VALUE :: flags
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_set_kpsi(h, kpsi) BIND(C, NAME='vk_gpe_set_kpsi')
      IMPORT :: C_INT, C_PTR, C_DOUBLE_COMPLEX
      TYPE(C_PTR), VALUE :: h
      COMPLEX(C_DOUBLE_COMPLEX) :: kpsi(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_fft2d(h, dir, xin, xout) BIND(C, NAME='vk_fft2d')
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: dir
      TYPE(C_PTR), VALUE :: xin, xout
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_get_psi(h, psi, kpsi) BIND(C, NAME='vk_gpe_get_psi')
      IMPORT :: C_INT, C_PTR, C_DOUBLE_COMPLEX
      TYPE(C_PTR), VALUE :: h
      COMPLEX(C_DOUBLE_COMPLEX) :: psi(*), kpsi(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_rk4_steps(h, n) BIND(C, NAME='vk_gpe_rk4_steps')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: n
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_argle_steps(h, n, res) BIND(C, NAME='vk_gpe_argle_steps')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: n
      REAL(C_DOUBLE) :: res
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_diagnostics(h, scal, spec, pscal) BIND(C, NAME='vk_gpe_diagnostics')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: scal(8), spec(*), pscal(5)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_ic_vortices_host(N, nv, xs, ys, qs, xi, nimg, psi) BIND(C, NAME='vk_ic_vortices_host')
      IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
      INTEGER(C_INT), VALUE :: N, nv, nimg
      REAL(C_DOUBLE) :: xs(*), ys(*)
      INTEGER(C_INT) :: qs(*)
      REAL(C_DOUBLE), VALUE :: xi
      COMPLEX(C_DOUBLE_COMPLEX) :: psi(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_wf_write(dir, iopoint, time, dt, N, nsplit, psi) BIND(C, NAME='vk_wf_write')
      IMPORT :: C_INT, C_DOUBLE, C_CHAR, C_DOUBLE_COMPLEX
      CHARACTER(KIND=C_CHAR) :: dir(*)
      INTEGER(C_INT), VALUE :: iopoint, N, nsplit
      REAL(C_DOUBLE), VALUE :: time, dt
      COMPLEX(C_DOUBLE_COMPLEX) :: psi(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_wf_read(dir, nrestart, N, nsplit_orig, iopoint, time, dt, psi) BIND(C, NAME='vk_wf_read')
      IMPORT :: C_INT, C_DOUBLE, C_CHAR, C_DOUBLE_COMPLEX
      CHARACTER(KIND=C_CHAR) :: dir(*)
      INTEGER(C_INT), VALUE :: nrestart, N, nsplit_orig
      INTEGER(C_INT) :: iopoint
      REAL(C_DOUBLE) :: time, dt
      COMPLEX(C_DOUBLE_COMPLEX) :: psi(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_set_advection(h, tg, vnx, vny) BIND(C, NAME='vk_gpe_set_advection')
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: tg
      REAL(C_DOUBLE), VALUE :: vnx, vny
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_sts_create(s, h, nt) BIND(C, NAME='vk_sts_create')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR) :: s
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: nt
    END FUNCTION
    SUBROUTINE vk_sts_destroy(s) BIND(C, NAME='vk_sts_destroy')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: s
    END SUBROUTINE
    INTEGER(C_INT) FUNCTION vk_sts_record(s) BIND(C, NAME='vk_sts_record')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_sts_shells(s) BIND(C, NAME='vk_sts_shells')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_sts_compute(s, spec2d, spec1d) BIND(C, NAME='vk_sts_compute')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: s
      REAL(C_DOUBLE) :: spec2d(*), spec1d(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_particles_set(h, pos, vel, fext) BIND(C, NAME='vk_gpe_particles_set')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: pos(*), vel(*), fext(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_particles_get(h, pos, vel, frc) BIND(C, NAME='vk_gpe_particles_get')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: pos(*), vel(*), frc(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_part_checkpoint(dir, nrst, iopoint, ptime, np, pos, vel, fext) &
        BIND(C, NAME='vk_part_checkpoint')
      IMPORT :: C_INT, C_DOUBLE, C_CHAR
      CHARACTER(KIND=C_CHAR) :: dir(*)
      INTEGER(C_INT), VALUE :: nrst, iopoint, np
      REAL(C_DOUBLE), VALUE :: ptime
      REAL(C_DOUBLE) :: pos(*), vel(*), fext(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_part_restart(dir, nrst, np, pos, vel, fext) BIND(C, NAME='vk_part_restart')
      IMPORT :: C_INT, C_DOUBLE, C_CHAR
      CHARACTER(KIND=C_CHAR) :: dir(*)
      INTEGER(C_INT), VALUE :: nrst, np
      REAL(C_DOUBLE) :: pos(*), vel(*), fext(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_set_sgle(h, on, oneybeta, nuN, chempot0, ktrunc, seed) BIND(C, NAME='vk_gpe_set_sgle')
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_LONG_LONG
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: on
      REAL(C_DOUBLE), VALUE :: oneybeta, nuN, chempot0, ktrunc
      INTEGER(C_LONG_LONG), VALUE :: seed
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_get_chempot(h, omega) BIND(C, NAME='vk_gpe_get_chempot')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE) :: omega
    END FUNCTION
    INTEGER(C_INT) FUNCTION vk_gpe_set_vortex_advection(h, nv, xs, ys, qs, nimg) BIND(C, NAME='vk_gpe_set_vortex_advection')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT), VALUE :: nv, nimg
      REAL(C_DOUBLE) :: xs(*), ys(*)
      INTEGER(C_INT) :: qs(*)
    END FUNCTION
    FUNCTION vk_last_error() BIND(C, NAME='vk_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR) :: vk_last_error
    END FUNCTION
  END INTERFACE
END MODULE vikshobha_c


MODULE vikshobha_gpu
  USE ISO_C_BINDING
  USE vikshobha_c
  IMPLICIT NONE

  INTEGER, PARAMETER :: forward = -1, inverse = 1      ! src/mod_fft.f90:48
  INTEGER :: Nx, Ny, nshell
  TYPE(vk_gpe_params) :: params
  TYPE(C_PTR) :: vkh = C_NULL_PTR
  COMPLEX(C_DOUBLE_COMPLEX), ALLOCATABLE, TARGET :: psi(:,:), kpsi(:,:)
  ! the reference's other transform work arrays (src/mod_fft.f90:77-118):
  ! each plan's physical (Nx,Ny) and Fourier (Ny,Nx, transposed) arrays share
  ! one buffer, as FFTW's in-place plans there do
  COMPLEX(C_DOUBLE_COMPLEX), ALLOCATABLE, TARGET :: work_store(:,:)
  COMPLEX(C_DOUBLE_COMPLEX), POINTER :: cgp_phys_nlin(:,:) => NULL(), cgp_four_nlin(:,:) => NULL()
  COMPLEX(C_DOUBLE_COMPLEX), POINTER :: phys_cmpxtmp(:,:) => NULL(), four_cmpxtmp(:,:) => NULL()
  COMPLEX(C_DOUBLE_COMPLEX), POINTER :: cgp_phys_tmp1(:,:) => NULL(), cgp_four_tmp1(:,:) => NULL()
  COMPLEX(C_DOUBLE_COMPLEX), POINTER :: cgp_phys_tmp2(:,:) => NULL(), cgp_four_tmp2(:,:) => NULL()
  COMPLEX(C_DOUBLE_COMPLEX), POINTER :: cgp_phys_tmp3(:,:) => NULL(), cgp_four_tmp3(:,:) => NULL()
  COMPLEX(C_DOUBLE_COMPLEX), POINTER :: cgp_phys_trans_vobj(:,:) => NULL(), cgp_four_trans_vobj(:,:) => NULL()
  REAL(C_DOUBLE) :: egrad, echem, eintr, etot, norm, momx, momy, n0, argle_residual
  REAL(C_DOUBLE), ALLOCATABLE :: spec_gp(:,:)
  ! checkpoint layout: slabs written per file (nsplit_current) and slabs the
  ! restart files were written with (nsplit_orignal); directories as in the
  ! reference driver
  INTEGER :: nsplit_orignal = 1, nsplit_current = 1
  CHARACTER(LEN=256) :: wf_out_dir = 'data/wf', wf_in_dir = 'restart/wf'
  REAL(C_DOUBLE) :: timestart = 0.0_C_DOUBLE, dt_read = 0.0_C_DOUBLE
  ! spatio-temporal spectra: Ntt snapshots (the CompSTS driver's Nt = Ntt),
  ! spec_sts2d(0:shells-1, 1:Ntt) and arr1(1:Ny, 1:Ntt) as CompSTS computes them
  INTEGER :: Ntt = 2048
  TYPE(C_PTR) :: vks = C_NULL_PTR
  REAL(C_DOUBLE), ALLOCATABLE :: spec_sts2d(:,:), arr1(:,:)
  ! the namelist inputs the argument-less initialize_fft reads
  ! (src/driver.f90:52-78; xi in units of dx, ppotV0 in units of g, ppotd in
  ! units of xi), the restart number that names the output files, and the
  ! diagnostic strides of src/diagnostics.f90:430-431
  INTEGER :: Nx_in = 0, nrst = 0, gsteps = 1, ssteps = 1
  REAL(C_DOUBLE) :: dt_fix = 0.0_C_DOUBLE, xi = 0.0_C_DOUBLE, csound = 0.0_C_DOUBLE
  REAL(C_DOUBLE) :: rho_avg = 1.0_C_DOUBLE, disp_factor = 0.0_C_DOUBLE
  CHARACTER(LEN=256) :: data_dir = 'data', spectra_dir = 'spectra', part_in_dir = 'restart/part'
  ! particles (mod_gpparticles, src/mod_gpparticles.f90:14-30): (ndim, gp_N_obj)
  INTEGER :: gp_N_obj = 0
  REAL(C_DOUBLE) :: ppotV0 = 0.0_C_DOUBLE, ppotd = 0.0_C_DOUBLE, gp_mass_part = 1.0_C_DOUBLE
  REAL(C_DOUBLE), ALLOCATABLE :: gp_pos_obj(:,:), gp_vel_obj(:,:), gp_f0ext_obj(:,:), gp_frc_obj(:,:)
  REAL(C_DOUBLE) :: eppot = 0.0_C_DOUBLE, eobj_kin = 0.0_C_DOUBLE, esrep = 0.0_C_DOUBLE, mom_obj(2) = 0.0_C_DOUBLE

  ! SGLE (GPE_SGLE, namelist gpfiniteT, src/driver.f90:59): oneybeta, nuN,
  ! chempot0, and GPE_SELFTRUNC's ktrunc; omega_chem follows the chemical
  ! potential the steps reach
  REAL(C_DOUBLE) :: oneybeta = 0.0_C_DOUBLE, nuN = 0.0_C_DOUBLE, chempot0 = 0.0_C_DOUBLE
  REAL(C_DOUBLE) :: ktrunc = 0.0_C_DOUBLE, omega_chem = 0.0_C_DOUBLE
  INTEGER(C_LONG_LONG) :: sgle_seed = 0_C_LONG_LONG

  INTERFACE initialize_fft
    MODULE PROCEDURE initialize_fft_namelist, initialize_fft_args
  END INTERFACE

CONTAINS

  SUBROUTINE vk_check(rc, where)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(*), INTENT(IN) :: where
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    INTEGER :: i
    IF (rc /= 0) THEN
      CALL C_F_POINTER(vk_last_error(), msg, [512])
      WRITE(*, '(A,A,I4,A)', ADVANCE='NO') where, ': vikshobha error ', rc, ': '
      DO i = 1, 512
        IF (msg(i) == C_NULL_CHAR) EXIT
        WRITE(*, '(A)', ADVANCE='NO') msg(i)
      END DO
      WRITE(*, *)
      STOP 2
    END IF
  END SUBROUTINE

  ! set_grid + constants (src/mod_grid.f90:63-101, src/driver.f90:360-376)
  ! followed by create_state_variables + initialize_fft, on the namelist
  ! globals Nx_in, dt_fix, xi, csound, rho_avg, disp_factor and, for
  ! gp_N_obj > 0, the particle inputs (ppotV0 = ppotV0*gstrength and
  ! ppotd = ppotd*xi as src/driver.f90:390-392; create_particles)
  SUBROUTINE initialize_fft_namelist()
    CALL vk_check(vk_gpe_params_init(params, INT(Nx_in, C_INT), dt_fix, xi, csound, rho_avg, disp_factor), &
                  'params_init')
    IF (gp_N_obj > 0) THEN
      params%np = INT(gp_N_obj, C_INT)
      params%V0 = ppotV0 * params%g
      params%d = ppotd * params%xi
      params%mass = gp_mass_part
      params%force_dealias = 1_C_INT
      ALLOCATE(gp_pos_obj(2, gp_N_obj), gp_vel_obj(2, gp_N_obj), gp_f0ext_obj(2, gp_N_obj), &
               gp_frc_obj(2, gp_N_obj))
      gp_pos_obj = 0.0_C_DOUBLE; gp_vel_obj = 0.0_C_DOUBLE
      gp_f0ext_obj = 0.0_C_DOUBLE; gp_frc_obj = 0.0_C_DOUBLE
    END IF
    CALL create_state()
  END SUBROUTINE

  SUBROUTINE initialize_fft_args(N, dt_in, xi_in, csound_in, rho_avg_in, disp_in)
    INTEGER, INTENT(IN) :: N
    REAL(C_DOUBLE), INTENT(IN) :: dt_in, xi_in, csound_in, rho_avg_in, disp_in
    Nx_in = N; dt_fix = dt_in; xi = xi_in; csound = csound_in; rho_avg = rho_avg_in; disp_factor = disp_in
    gp_N_obj = 0
    CALL initialize_fft_namelist()
  END SUBROUTINE

  SUBROUTINE create_state()
    CALL vk_check(vk_gpe_create(vkh, params), 'create')
    Nx = Nx_in; Ny = Nx_in
    nshell = vk_nshell(INT(Nx, C_INT))
    ALLOCATE(psi(Nx, Ny), kpsi(Ny, Nx), spec_gp(0:nshell, 1:6))
    psi = (0.0_C_DOUBLE, 0.0_C_DOUBLE); kpsi = psi
    ALLOCATE(work_store(Nx * Ny, 6))
    work_store = (0.0_C_DOUBLE, 0.0_C_DOUBLE)
    CALL C_F_POINTER(C_LOC(work_store(1, 1)), cgp_phys_nlin, [Nx, Ny])
    CALL C_F_POINTER(C_LOC(work_store(1, 1)), cgp_four_nlin, [Ny, Nx])
    CALL C_F_POINTER(C_LOC(work_store(1, 2)), phys_cmpxtmp, [Nx, Ny])
    CALL C_F_POINTER(C_LOC(work_store(1, 2)), four_cmpxtmp, [Ny, Nx])
    CALL C_F_POINTER(C_LOC(work_store(1, 3)), cgp_phys_tmp1, [Nx, Ny])
    CALL C_F_POINTER(C_LOC(work_store(1, 3)), cgp_four_tmp1, [Ny, Nx])
    CALL C_F_POINTER(C_LOC(work_store(1, 4)), cgp_phys_tmp2, [Nx, Ny])
    CALL C_F_POINTER(C_LOC(work_store(1, 4)), cgp_four_tmp2, [Ny, Nx])
    CALL C_F_POINTER(C_LOC(work_store(1, 5)), cgp_phys_tmp3, [Nx, Ny])
    CALL C_F_POINTER(C_LOC(work_store(1, 5)), cgp_four_tmp3, [Ny, Nx])
    CALL C_F_POINTER(C_LOC(work_store(1, 6)), cgp_phys_trans_vobj, [Nx, Ny])
    CALL C_F_POINTER(C_LOC(work_store(1, 6)), cgp_four_trans_vobj, [Ny, Nx])
  END SUBROUTINE

  FUNCTION fnrst() RESULT(s)
    CHARACTER(LEN=16) :: s
    WRITE(s, '(I8)') nrst
    s = ADJUSTL(s)
  END FUNCTION

  ! fft(plan_name, direction) (src/mod_fft.f90:471-542).  Plan 'psi' acts
  ! on the solver state (forward: kpsi = FFT(psi); inverse: psi =
  ! IFFT(kpsi)/(Nx*Ny)) and refreshes the host copies.  The work-array plans
  ! 'gpnlin', 'cmpxtmp', 'gptmp1'..'gptmp3' and 'gptransvobj' transform their
  ! shared buffer in place on the GPU: forward takes the physical array to
  ! the transposed Fourier array, unnormalised; inverse scales by 1/(Nx*Ny).
  SUBROUTINE fft(plan_name, direction)
    CHARACTER(*), INTENT(IN) :: plan_name
    INTEGER, INTENT(IN) :: direction
    INTEGER :: k
    IF (TRIM(plan_name) == 'psi') THEN
      IF (direction == forward) THEN
        CALL vk_check(vk_gpe_set_psi(vkh, psi, 0_C_INT), 'fft forward')
      ELSE
        CALL vk_check(vk_gpe_set_kpsi(vkh, kpsi), 'fft inverse')
      END IF
      CALL sync_host()
      RETURN
    END IF
    SELECT CASE (TRIM(plan_name))
    CASE ('gpnlin'); k = 1
    CASE ('cmpxtmp'); k = 2
    CASE ('gptmp1'); k = 3
    CASE ('gptmp2'); k = 4
    CASE ('gptmp3'); k = 5
    CASE ('gptransvobj'); k = 6
    CASE DEFAULT
      WRITE(*,*) 'fft: unknown plan ', TRIM(plan_name)
      STOP 2
    END SELECT
    CALL vk_check(vk_fft2d(vkh, INT(SIGN(1, direction), C_INT), C_LOC(work_store(1, k)), &
                           C_LOC(work_store(1, k))), 'fft '//TRIM(plan_name))
  END SUBROUTINE

  ! restart_from_argle_data + ic_otoc_gpe: kpsi = FFT(psi), dealias, psi = IFFT(kpsi)
  SUBROUTINE set_initial_psi()
    CALL vk_check(vk_gpe_set_psi(vkh, psi, 1_C_INT), 'set_initial_psi')
    CALL sync_host()
  END SUBROUTINE

  SUBROUTINE sync_host()
    CALL vk_check(vk_gpe_get_psi(vkh, psi, kpsi), 'sync_host')
  END SUBROUTINE

  SUBROUTINE gpe_rk4()
    CALL vk_check(vk_gpe_rk4_steps(vkh, 1_C_INT), 'gpe_rk4')
  END SUBROUTINE

  SUBROUTINE gpe_rk4_steps(n)
    INTEGER, INTENT(IN) :: n
    CALL vk_check(vk_gpe_rk4_steps(vkh, INT(n, C_INT)), 'gpe_rk4_steps')
  END SUBROUTINE

  SUBROUTINE ARGLE(time)
    REAL(C_DOUBLE), INTENT(IN) :: time
    CALL vk_check(vk_gpe_argle_steps(vkh, 1_C_INT, argle_residual), 'ARGLE')
    CALL vk_check(vk_gpe_get_chempot(vkh, omega_chem), 'ARGLE')
  END SUBROUTINE

  ! the GPE_SGLE build of ARGLE on the gpfiniteT globals (noise from the
  ! library's generator, seeded by sgle_seed)
  SUBROUTINE set_sgle(on)
    LOGICAL, INTENT(IN) :: on
    INTEGER(C_INT) :: o
    o = 0
    IF (on) o = 1
    CALL vk_check(vk_gpe_set_sgle(vkh, o, oneybeta, nuN, chempot0, ktrunc, sgle_seed), 'set_sgle')
    omega_chem = chempot0
  END SUBROUTINE

  ! ARGLE advected by the imposed velocity of a vortex configuration (the
  ! current of its Pade product ansatz), which keeps the vortices in place
  SUBROUTINE set_vortex_advection(nv, xs, ys, qs)
    INTEGER, INTENT(IN) :: nv
    REAL(C_DOUBLE), INTENT(IN) :: xs(nv), ys(nv)
    INTEGER(C_INT), INTENT(IN) :: qs(nv)
    CALL vk_check(vk_gpe_set_vortex_advection(vkh, INT(nv, C_INT), xs, ys, qs, 3_C_INT), 'set_vortex_advection')
  END SUBROUTINE

  ! diagnostics_gpe (src/diagnostics.f90:4-571): energies, norm, momentum and
  ! the six spectra, appended to data/energy1, norm and mom.dat_nrst<nrst>
  ! (:240-259); with particles the forces and the per-particle and energy
  ! files (:523-569); quantglobal_gpe every gsteps and spectrum_gpe every
  ! ssteps (:430-431)
  SUBROUTINE diagnostics_gpe(spoint, jtime, time)
    INTEGER, INTENT(IN) :: spoint, jtime
    REAL(C_DOUBLE), INTENT(IN) :: time
    REAL(C_DOUBLE) :: scal(8), pscal(5)
    INTEGER :: u, i1
    CHARACTER(LEN=16) :: nobj
    CALL vk_check(vk_gpe_diagnostics(vkh, scal, spec_gp, pscal), 'diagnostics_gpe')
    egrad = scal(1); echem = scal(2); eintr = scal(3); etot = scal(4)
    norm = scal(5); momx = scal(6); momy = scal(7); n0 = scal(8)
    OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/energy1.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
    WRITE(u, '(E15.6,E15.6,E15.6,E15.6,E15.6)') time, egrad, echem, eintr, etot
    CLOSE(u)
    OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/norm.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
    WRITE(u, '(E15.6,E15.6)') time, norm
    CLOSE(u)
    OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/mom.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
    WRITE(u, '(E15.6,E15.6,E15.6)') time, momx, momy
    CLOSE(u)
    IF (MOD(jtime, gsteps) == 0) CALL quantglobal_gpe(jtime, time)
    IF (MOD(jtime, ssteps) == 0) CALL spectrum_gpe(spoint, time)
    IF (params%np > 0) THEN
      eppot = pscal(1); eobj_kin = pscal(2); esrep = pscal(3); mom_obj = pscal(4:5)
      CALL vk_check(vk_gpe_particles_get(vkh, gp_pos_obj, gp_vel_obj, gp_frc_obj), 'diagnostics_gpe')
      DO i1 = 1, gp_N_obj
        WRITE(nobj, '(I8)') i1
        nobj = ADJUSTL(nobj)
        OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/gpobj_pos_'//TRIM(nobj)//'.dat_nrst'//TRIM(fnrst()), &
             POSITION='APPEND')
        WRITE(u, '(E15.6,E15.6,E15.6)') time, gp_pos_obj(1, i1), gp_pos_obj(2, i1)
        CLOSE(u)
        OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/gpobj_vel_'//TRIM(nobj)//'.dat_nrst'//TRIM(fnrst()), &
             POSITION='APPEND')
        WRITE(u, '(E15.6,E15.6,E15.6)') time, gp_vel_obj(1, i1), gp_vel_obj(2, i1)
        CLOSE(u)
        OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/gpobj_frc_'//TRIM(nobj)//'.dat_nrst'//TRIM(fnrst()), &
             POSITION='APPEND')
        WRITE(u, '(E15.6,E15.6,E15.6)') time, gp_frc_obj(1, i1), gp_frc_obj(2, i1)
        CLOSE(u)
      END DO
      OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/egpobj.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
      WRITE(u, '(E15.6,E15.6,E15.6,E15.6)') time, eppot, eobj_kin, esrep
      CLOSE(u)
      OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/momgpobj.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
      WRITE(u, '(E15.6,E15.6,E15.6)') time, mom_obj(1), mom_obj(2)
      CLOSE(u)
    END IF
  END SUBROUTINE

  ! quantglobal_gpe (src/diagnostics.f90:573-606): spectral totals and the
  ! condensate occupation n0
  SUBROUTINE quantglobal_gpe(jtime, time)
    INTEGER, INTENT(IN) :: jtime
    REAL(C_DOUBLE), INTENT(IN) :: time
    INTEGER :: u
    OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/energy.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
    WRITE(u, '(E15.6,E15.6,E15.6,E15.6,E15.6,E15.6)') time, SUM(spec_gp(1:nshell, 1)), &
        SUM(spec_gp(1:nshell, 2)), SUM(spec_gp(1:nshell, 3)), SUM(spec_gp(1:nshell, 4)), SUM(spec_gp(0:nshell, 5))
    CLOSE(u)
    OPEN(NEWUNIT=u, FILE=TRIM(data_dir)//'/n0.dat_nrst'//TRIM(fnrst()), POSITION='APPEND')
    WRITE(u, '(E15.6,E15.6)') time, spec_gp(0, 6)
    CLOSE(u)
  END SUBROUTINE

  ! spectrum_gpe (src/diagnostics.f90:608-634): spectra/spec<spoint>.dat,
  ! shells 1..nshell
  SUBROUTINE spectrum_gpe(spoint, time)
    INTEGER, INTENT(IN) :: spoint
    REAL(C_DOUBLE), INTENT(IN) :: time
    INTEGER :: u, i1
    CHARACTER(LEN=16) :: fnum
    WRITE(fnum, '(I8)') spoint
    fnum = ADJUSTL(fnum)
    OPEN(NEWUNIT=u, FILE=TRIM(spectra_dir)//'/spec'//TRIM(fnum)//'.dat', STATUS='UNKNOWN')
    DO i1 = 1, nshell
      WRITE(u, '(E15.6,E15.6,E15.6,E15.6,E15.6,E15.6,E15.6)') DBLE(i1), spec_gp(i1, 1), spec_gp(i1, 2), &
          spec_gp(i1, 3), spec_gp(i1, 4), spec_gp(i1, 5), spec_gp(i1, 6)
    END DO
    CLOSE(u)
  END SUBROUTINE

  ! particle_potential (src/gpe_particle.f90:3-90): the potential of one
  ! particle at the origin, transformed and dealiased on the device (the
  ! particle state is handed over with it)
  SUBROUTINE particle_potential()
    CALL vk_check(vk_gpe_particles_set(vkh, gp_pos_obj, gp_vel_obj, gp_f0ext_obj), 'particle_potential')
  END SUBROUTINE

  ! initialize_particles (src/gpe_particle.f90:92-141): positions,
  ! velocities and external forces from
  ! <part_in_dir>/gppos|gpvel|gpf0ext.dat_nrst<nrst>, then to the device
  SUBROUTINE initialize_particles()
    CALL vk_check(vk_part_restart(TRIM(part_in_dir)//C_NULL_CHAR, INT(nrst, C_INT), INT(gp_N_obj, C_INT), &
                                  gp_pos_obj, gp_vel_obj, gp_f0ext_obj), 'initialize_particles')
    CALL vk_check(vk_gpe_particles_set(vkh, gp_pos_obj, gp_vel_obj, gp_f0ext_obj), 'initialize_particles')
  END SUBROUTINE

  ! gppart_checkpoint (src/gpe_particle.f90:143-184): appends the current
  ! positions, velocities and external forces to
  ! <data_dir>/gppos|gpvel|gpf0ext_ckp.dat_nrst<nrst>
  SUBROUTINE gppart_checkpoint(iopoint, jtime, ptime)
    INTEGER, INTENT(IN) :: iopoint, jtime
    REAL(C_DOUBLE), INTENT(IN) :: ptime
    CALL vk_check(vk_gpe_particles_get(vkh, gp_pos_obj, gp_vel_obj, gp_frc_obj), 'gppart_checkpoint')
    CALL vk_check(vk_part_checkpoint(TRIM(data_dir)//C_NULL_CHAR, INT(nrst, C_INT), INT(iopoint, C_INT), ptime, &
                                     INT(gp_N_obj, C_INT), gp_pos_obj, gp_vel_obj, gp_f0ext_obj), 'gppart_checkpoint')
  END SUBROUTINE

  ! io_fields (src/io_fields.f90:41-50): psi to wf<iopoint>p<r>.dat
  SUBROUTINE io_fields(iopoint, time)
    INTEGER, INTENT(IN) :: iopoint
    REAL(C_DOUBLE), INTENT(IN) :: time
    CALL sync_host()
    CALL vk_check(vk_wf_write(TRIM(wf_out_dir)//C_NULL_CHAR, INT(iopoint, C_INT), time, params%dt, &
                              INT(Nx, C_INT), INT(nsplit_current, C_INT), psi), 'io_fields')
  END SUBROUTINE

  ! restart_gpe (src/io_fields.f90:71-224): psi from the nsplit_orignal wf
  ! files of checkpoint nrestart, then kpsi = FFT(psi)
  SUBROUTINE restart_gpe(nrestart)
    INTEGER, INTENT(IN) :: nrestart
    INTEGER(C_INT) :: icpoint
    CALL vk_check(vk_wf_read(TRIM(wf_in_dir)//C_NULL_CHAR, INT(nrestart, C_INT), INT(Nx, C_INT), &
                             INT(nsplit_orignal, C_INT), icpoint, timestart, dt_read, psi), 'restart_gpe')
    CALL fft('psi', forward)
  END SUBROUTINE

  ! the GPE_ADVECTION / GPE_COUNTERFLOW builds of ARGLE (src/gpe.f90:78-146):
  ! taylor_green selects the Taylor-Green term, (vnx, vny) the counterflow
  SUBROUTINE set_argle_advection(taylor_green, vnx, vny)
    LOGICAL, INTENT(IN) :: taylor_green
    REAL(C_DOUBLE), INTENT(IN) :: vnx, vny
    INTEGER(C_INT) :: tg
    tg = 0
    IF (taylor_green) tg = 1
    CALL vk_check(vk_gpe_set_advection(vkh, tg, vnx, vny), 'set_argle_advection')
  END SUBROUTINE

  ! sts_gpe (src/diagnostics.f90:636-695): one snapshot of kpsi, kept on the
  ! device (the first call allocates Ntt snapshots)
  SUBROUTINE sts_gpe(stspoint, time)
    INTEGER, INTENT(IN) :: stspoint
    REAL(C_DOUBLE), INTENT(IN) :: time
    IF (.NOT. C_ASSOCIATED(vks)) THEN
      CALL vk_check(vk_sts_create(vks, vkh, INT(Ntt, C_INT)), 'sts_gpe')
      ALLOCATE(spec_sts2d(0:vk_sts_shells(vks) - 1, 1:Ntt), arr1(1:Ny, 1:Ntt))
    END IF
    CALL vk_check(vk_sts_record(vks), 'sts_gpe')
  END SUBROUTINE

  ! the CompSTS post-processing (Run/CompSTS/driver.f90:134-265) once Ntt
  ! snapshots are taken: fills spec_sts2d and arr1
  SUBROUTINE comp_sts()
    CALL vk_check(vk_sts_compute(vks, spec_sts2d, arr1), 'comp_sts')
  END SUBROUTINE

  SUBROUTINE clear_fft()
    IF (C_ASSOCIATED(vks)) CALL vk_sts_destroy(vks)
    vks = C_NULL_PTR
    IF (ALLOCATED(spec_sts2d)) DEALLOCATE(spec_sts2d, arr1)
    IF (C_ASSOCIATED(vkh)) CALL vk_gpe_destroy(vkh)
    vkh = C_NULL_PTR
    IF (ALLOCATED(psi)) DEALLOCATE(psi, kpsi, spec_gp)
    IF (ALLOCATED(work_store)) THEN
      NULLIFY(cgp_phys_nlin, cgp_four_nlin, phys_cmpxtmp, four_cmpxtmp, cgp_phys_tmp1, cgp_four_tmp1, &
              cgp_phys_tmp2, cgp_four_tmp2, cgp_phys_tmp3, cgp_four_tmp3, cgp_phys_trans_vobj, cgp_four_trans_vobj)
      DEALLOCATE(work_store)
    END IF
    IF (ALLOCATED(gp_pos_obj)) DEALLOCATE(gp_pos_obj, gp_vel_obj, gp_f0ext_obj, gp_frc_obj)
  END SUBROUTINE

END MODULE vikshobha_gpu
