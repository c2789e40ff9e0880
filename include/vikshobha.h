/*
 * vikshobha: MI355X (gfx950) implementation of the 2D Gross-Pitaevskii
 * pseudospectral step of SourangshuGhosh/QuantumVortices ("VIKSHOBHA").
 *
 * C ABI.  Plain pointers and sizes; complex numbers are interleaved
 * (re, im) doubles.  Host arrays use the reference's Fortran views:
 *   psi(i1=x, i2=y), x fastest            -> psi_xy[2*(y*N + x) + {0,1}]
 *   kpsi(i2=ky, i1=kx), ky fastest        -> kpsi_t[2*(kx*N + ky) + {0,1}]
 *     (FFTW_MPI_TRANSPOSED_OUT layout, src/mod_fft.f90:143-146,307-311)
 *
 * Return values: 0 = success, < 0 = error (see vk_last_error(), which is
 * thread-local).  One handle per host thread; each handle owns its device
 * buffers and one HIP stream; calls are synchronous unless named _async.
 * Every compute entry point requires a ROCm device; there is no CPU path.
 */
#ifndef VIKSHOBHA_H
#define VIKSHOBHA_H

#ifdef __cplusplus
extern "C" {
#endif

#define VK_OK 0
#define VK_ERR_ARG -1
#define VK_ERR_HIP -2
#define VK_ERR_STATE -3
#define VK_ERR_UNSUPPORTED -4

/* set_psi flags */
#define VK_SYNC_DEALIAS 1  /* after kpsi = FFT(psi): dealias kpsi and
                              psi = IFFT(kpsi), as restart_from_argle_data +
                              ic_otoc_gpe do (src/initial_config.f90:759,
                              :316-343) */

typedef struct vk_gpe vk_gpe;

/* Run parameters (the reference's module globals after src/driver.f90:360-398
 * and src/mod_grid.f90:63-101). */
typedef struct {
    int N;              /* Nx = Ny, power of two, 32..4096 */
    double dt;          /* dt_fix */
    double alpha;       /* alpha */
    double g;           /* gstrength */
    double omega;       /* omega_chem */
    double disp;        /* disp_factor */
    double xi;          /* xi (already multiplied by dx) */
    /* particles (GPPART): number, ppotV0 (already x g), ppotd (already x xi),
     * gp_mass_part, and 1 to dealias rho-hat in the particle force with
     * k1^2+k2^2 (the reference's intent; src/gpe_particle.f90:275 tests an
     * uninitialised ksqr) */
    int np;
    double V0, d, mass;
    int force_dealias;
} vk_gpe_params;

/* Grid constants and derived GPE constants in the reference's order of
 * operations: src/mod_grid.f90:63-101 (dx, kalias, nshell) and
 * src/driver.f90:360-376 (xi*=dx, alpha, gstrength, omega_chem). */
int vk_gpe_params_init(vk_gpe_params* p, int N, double dt, double xi_in,
                       double csound, double rho_avg, double disp);
int vk_nshell(int N);               /* int(1.414*Nx/2.0d0)+1 */

/* Replaces create_state_variables + initialize_fft (src/mod_fft.f90:122-198,
 * :288-428; FFTW plans + fftw_alloc_complex) and global_array
 * (src/pspectral.f90:1-35). */
int vk_gpe_create(vk_gpe** h, const vk_gpe_params* p);
void vk_gpe_destroy(vk_gpe* h);    /* clear_fft + free_state_variables */

/* Field upload/download.  set_psi: kpsi = FFT(psi) (src/initial_config.f90:759)
 * and, with VK_SYNC_DEALIAS, the dealias + psi=IFFT(kpsi) sync of
 * ic_otoc_gpe.  get_psi: either pointer may be NULL.  The arrays may be host
 * or device pointers; device data must be complete when the call is made
 * (the copies run on the handle's stream, which does not wait for other
 * streams), and each call returns after its copies have finished. */
int vk_gpe_set_psi(vk_gpe* h, const double* psi_xy, int flags);
int vk_gpe_set_kpsi(vk_gpe* h, const double* kpsi_t);
int vk_gpe_get_psi(vk_gpe* h, double* psi_xy, double* kpsi_t);

/* nsteps x gpe_rk4 (src/gpe.f90:310-494 with gpe_rhs :498-671; with
 * particles also gp_particle_motion_rk4, src/gpe_particle.f90:429-531). */
int vk_gpe_rk4_steps(vk_gpe* h, int nsteps);
int vk_gpe_rk4_steps_async(vk_gpe* h, int nsteps);

/* nsteps x ARGLE(time) (src/gpe.f90:5-224, plain variant).  residual (may be
 * NULL) receives ||kpsi_new-kpsi_old||_2/(dt*||kpsi_old||_2) of the last
 * step; the reference computes no residual. */
int vk_gpe_argle_steps(vk_gpe* h, int nsteps, double* residual);
/* ARGLE until the residual (as above, evaluated every check_every steps)
 * is <= tol or max_steps are done; steps_out receives the steps taken. */
int vk_gpe_argle_until(vk_gpe* h, double tol, int max_steps, int check_every,
                       int* steps_out, double* residual);

/* diagnostics_gpe energies and spectra (src/diagnostics.f90:54-426):
 * scal[8] = {egrad, echem, eintr, etot, norm, momx, momy, n0=spec(0,6)};
 * spec = spec_gp(0:nshell,1:6) in Fortran order (may be NULL);
 * pscal[5] = {eppot, eobj_kin, esrep, momobj_x, momobj_y} (particles; may be
 * NULL). */
int vk_gpe_diagnostics(vk_gpe* h, double* scal, double* spec, double* pscal);

/* Particles: particle_potential (src/gpe_particle.f90:3-90) runs inside
 * particles_set; pos/vel/fext are (2,np) Fortran arrays (x,y per particle).
 * particles_get: frc receives the forces of the last diagnostics call
 * (gp_frc_obj, src/diagnostics.f90:528-533).  Any pointer may be NULL. */
int vk_gpe_particles_set(vk_gpe* h, const double* pos, const double* vel,
                         const double* fext);
int vk_gpe_particles_get(vk_gpe* h, double* pos, double* vel, double* frc);

/* Static obstacles for ARGLE (which never moves particles, src/gpe.f90:45-50):
 * replaces the particle count (at most the np the handle was created with),
 * positions pos[2*np], V0 and d, recomputes the potential and clears
 * velocities.  np = 0 removes the potential. */
int vk_gpe_set_obstacles(vk_gpe* h, int np, const double* pos, double V0, double d);

/* Replaces the time step (dt_fix) of a live handle. */
int vk_gpe_set_dt(vk_gpe* h, double dt);

/* Advective ARGLE (src/gpe.f90:78-81,91-94,114-146): taylor_green = 1 is
 * the GPE_ADVECTION build (u = (sin x cos y, -cos x sin y), terms
 * -i u.grad psi - |u|^2 psi/(4 alpha)); (vnx, vny) != 0 the GPE_COUNTERFLOW
 * build with namelist gpcounterflow (term -i v.grad psi).  Both may be on;
 * RK4 is unaffected, as in the reference. */
int vk_gpe_set_advection(vk_gpe* h, int taylor_green, double vnx, double vny);

/* Advective ARGLE with a caller-supplied velocity u(x,y) in place of the
 * Taylor-Green field of the GPE_ADVECTION slot (src/gpe.f90:114-130; same
 * terms -i u.grad psi - |u|^2 psi/(4 alpha), the counterflow of
 * vk_gpe_set_advection still added): ux, uy are (Nx,Ny) real arrays, host
 * or device; both NULL turns the field off.
 * vk_gpe_set_vortex_advection sets the imposed velocity of a vortex
 * configuration, the current of its Padé product ansatz
 * u = 2 alpha rho0 grad(theta0) (smooth, zero in the cores): ARGLE then
 * relaxes the cores in place instead of letting vortices drift and
 * annihilate.  vk_vortex_velocity_host computes the same field on the CPU. */
int vk_gpe_set_advection_field(vk_gpe* h, const double* ux, const double* uy);
int vk_gpe_set_vortex_advection(vk_gpe* h, int nv, const double* xs, const double* ys, const int* qs, int nimg);
int vk_vortex_velocity_host(int N, int nv, const double* xs, const double* ys, const int* qs, double xi,
                            double alpha, int nimg, double* ux, double* uy);

/* SGLE, the GPE_SGLE build of ARGLE(time) (src/gpe.f90:149-217 noise and
 * update, :226-266 finiteT_adhoc_chempot; namelist gpfiniteT
 * oneybeta, nuN, chempot0, src/driver.f90:59,376).  With on = 1 every ARGLE
 * step adds sqrt(dt) P_G[FFT(noisesigma xi/sqrt(dvol))], noisesigma =
 * sqrt(oneybeta/(2 alpha)), and first relaxes the chemical potential,
 * omega -= dt (nuN/L^2)(norm - L^2), starting from chempot0; ktrunc > 0
 * zeroes kpsi beyond |k| = ktrunc (GPE_SELFTRUNC).  xi is a field of
 * complex standard normals: vk_gpe_argle_steps / _until draw it from a
 * counter-based generator (Philox4x32-10 keyed by seed, draw n of the
 * handle's sequence, Box-Muller); the reference's Fortran RANDOM_NUMBER
 * stream cannot be reproduced, only its distribution.
 * vk_gpe_argle_step_noise runs one step with the caller's xi
 * (noise_xy: (Nx,Ny) complex, host or device).  vk_sgle_noise writes draw
 * `draw` of the generator (unscaled).  vk_gpe_get_chempot returns the
 * current omega. */
int vk_gpe_set_sgle(vk_gpe* h, int on, double oneybeta, double nuN, double chempot0, double ktrunc,
                    unsigned long long seed);
int vk_gpe_argle_step_noise(vk_gpe* h, const double* noise_xy, double* residual);
int vk_sgle_noise(vk_gpe* h, unsigned long long draw, double* out_xy);
int vk_gpe_get_chempot(vk_gpe* h, double* omega);

/* Conformance hook mirroring fft(plan_name, direction)
 * (src/mod_fft.f90:471-542): dir = -1: in = x-space (Nx,Ny) -> out = k-space
 * (Ny,Nx) transposed, unnormalised; dir = +1: inverse times 1/(Nx*Ny). */
int vk_fft2d(vk_gpe* h, int dir, const double* in, double* out);

/* Allocation-free variant of vk_fft2d (no counterpart in the reference; for
 * callers that cannot allocate: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  in and out are distinct device
 * buffers of N*N complex doubles in vk_fft2d's layouts; the intermediate
 * lives in a fixed library scratch of 1024^2 complex doubles (16 MiB per
 * GPU, held until the process ends), so N <= 1024 (larger N:
 * VK_ERR_UNSUPPORTED).  No hipMalloc, hipFree or other allocation happens
 * in the call.  Stream-ordered on the handle's stream (vk_gpe_stream): in
 * must be complete when the call is made, out is complete once that stream
 * has drained.  All calls on one GPU share the scratch: calls that overlap
 * on one GPU (other handles' streams, other host threads) race. */
int vk_fft2d_noalloc(vk_gpe* h, int dir, const double* in, double* out);

/* Padé-core vortex product initial condition (no reference counterpart:
 * the reference reads its initial state from restart files).
 * nv vortices at (xs, ys) with charges qs (+-1), healing length xi,
 * rho(r) = r^2 (a1 + a2 r^2)/(1 + b1 r^2 + a2 r^4), r in units of xi, and
 * the periodic image-sum phase; output psi_xy (Nx,Ny).  The _host variant
 * runs on the CPU; vk_gpe_set_ic_vortices builds the field on the device
 * and sets it as in vk_gpe_set_psi(h, ., flags). */
int vk_ic_vortices_host(int N, int nv, const double* xs, const double* ys,
                        const int* qs, double xi, int nimg, double* psi_xy);
int vk_gpe_set_ic_vortices(vk_gpe* h, int nv, const double* xs,
                           const double* ys, const int* qs, int nimg, int flags);

/* Timing support.  stream: the handle's hipStream_t.  pass_times: runs
 * nsteps RK4 steps with an event around every kernel launch and returns the
 * average duration (ms) per kernel kind in ms_out[nkinds] (names via
 * vk_pass_name); the state advances by nsteps. */
void* vk_gpe_stream(vk_gpe* h);
int vk_gpe_pass_times(vk_gpe* h, int nsteps, double* ms_out, int nkinds);
/* Row-pass variants of the psi-row path (N >= 2048; no reference
 * counterpart, the reference has one FFT plan per size).  p1: 0 TR rows per
 * workgroup, 1 one row per workgroup, 2 the persistent P1 that prefetches
 * its next rows into LDS (4096^2 on a device whose CU count divides the row
 * blocks evenly over the XCDs); p3: 0 TR rows, 1 one row.  The variants
 * order floating-point sums differently: results agree to roundoff, not
 * bitwise.  vk_gpe_autotune_rows times nsteps RK4 steps with every variant
 * pair from the current state, restores the state (kpsi, its y-inverse and
 * the particle state), keeps the fastest pair (choice[2] = {p1, p3}) and
 * returns ms/step in ms_out[2*p1 + p3] (6 entries, -1 where a variant does
 * not exist). */
/* Transform shape of a single field of N in {256, 512, 1024} without
 * particles, whose RK4 step runs the batched-ensemble kernels on one
 * member: 0 (default; bitwise the result of a default ensemble member) or 1
 * (8 points per thread, radix-8 stages, half-width panels: more, smaller
 * workgroups, faster on one field; results agree to roundoff).  No
 * reference counterpart. */
int vk_gpe_set_shape(vk_gpe* h, int var);
int vk_gpe_get_shape(vk_gpe* h, int* var);
int vk_gpe_set_row_variants(vk_gpe* h, int p1, int p3);
int vk_gpe_get_row_variants(vk_gpe* h, int* p1, int* p3);
int vk_gpe_autotune_rows(vk_gpe* h, int nsteps, int* choice, double* ms_out);
const char* vk_pass_name(int kind);
int vk_pass_kinds(void);
/* Compulsory HBM bytes one launch of kernel kind moves at this handle's N. */
double vk_gpe_pass_bytes(vk_gpe* h, int kind);
/* Measured attainable bandwidth (GB/s, read + written bytes per launch over
 * the average of reps launches) of the passes' access patterns on a 4096^2
 * complex128 field with no arithmetic: gbps[0] contiguous copy, [1] the
 * row passes' gather of 64-byte panel pieces, [2] the column pass copying
 * whole 2-column panels, [3] one column per workgroup (n <= 4 entries). */
int vk_pattern_bandwidth(int reps, double* gbps, int n);

/* Ensemble of independent realizations (no reference counterpart: the
 * reference runs one field per MPI job).  Members have N in {128, 256, 512,
 * 1024}, no particles; each pass of an RK4 step is one launch over all
 * members (mode 0), one per size (mode 1) or one per member (mode 2), or
 * the whole step is one persistent launch (mode 3: as many workgroups as
 * the device holds at once take the blocks of each pass in turn, with a
 * grid barrier between passes; a barrier that waits longer than one second
 * makes vk_ens_rk4_steps fail with VK_ERR_STATE and the state invalid), and
 * a step is replayed as a captured HIP graph (use_graph).  vk_ens_autotune
 * times every schedule on the current state, restores the state and keeps
 * the fastest (ms_out[8] = ms/step for modes 0..3 x {graph, direct}; -1 for
 * a mode 3 whose barriers timed out). */
typedef struct vk_ens vk_ens;
int vk_ens_create(vk_ens** e, int nreal, const vk_gpe_params* params);
void vk_ens_destroy(vk_ens* e);
int vk_ens_size(vk_ens* e);
void* vk_ens_stream(vk_ens* e);
int vk_ens_set_psi(vk_ens* e, int r, const double* psi_xy, int flags);
int vk_ens_set_ic_vortices(vk_ens* e, int r, int nv, const double* xs, const double* ys,
                           const int* qs, int nimg, int flags);
int vk_ens_get_psi(vk_ens* e, int r, double* psi_xy, double* kpsi_t);
int vk_ens_diagnostics(vk_ens* e, int r, double* scal, double* spec, double* pscal);
/* vk_ens_diagnostics of every member at once: the members' pipelines run
 * concurrently on their own streams with one wait at the end.  scal[8 r ..]
 * holds member r's scalars; member r's six spectra start at
 * spec[sum over q < r of 6 (nshell(N_q) + 1)] (the layout vk_gpe_diagnostics
 * gives one member). */
int vk_ens_diagnostics_all(vk_ens* e, double* scal, double* spec);
int vk_ens_rk4_steps(vk_ens* e, int nsteps);
int vk_ens_rk4_steps_async(vk_ens* e, int nsteps);
int vk_ens_set_mode(vk_ens* e, int mode, int use_graph);
int vk_ens_get_mode(vk_ens* e, int* mode, int* use_graph);
int vk_ens_autotune(vk_ens* e, int nsteps, double* ms_out);
/* Member shapes.  Sizes 256..1024 have two shapes with 256-thread
 * workgroups: 0 (16 points per thread, Cfg<N>) and 1 (8 points per thread,
 * radix-8 stages, half the panel width).  vk_ens_autotune_shapes times both
 * for each such size in the current launch mode (one size at a time,
 * ascending, kpsi restored before every trial), keeps the faster and
 * remembers it per (N, GPU model) for ensembles created later in the
 * process; vars_out[i] and ms_out[2i..2i+1] (ms/step of shapes 0, 1) are
 * per distinct member size, ascending (-1: one shape only).  Members of
 * shape 1 equal single-field runs to roundoff (the transforms factor
 * differently), members of shape 0 bit for bit. */
int vk_ens_set_shape(vk_ens* e, int N, int var);
int vk_ens_get_shape(vk_ens* e, int N, int* var);
int vk_ens_autotune_shapes(vk_ens* e, int nsteps, int* vars_out, double* ms_out);
/* drop the remembered shapes: later ensembles start in shape 0 */
void vk_ens_forget_shapes(void);

/* Fourier neural operator for the ARGLE warm start (no reference counterpart:
 * the reference starts ARGLE from the analytic ansatz of
 * src/initial_config.f90).  BF16 weights, FP32 activations and accumulation;
 * forward, backward and Adam run on the handle's stream.  Arrays are planar
 * FP32 [nb][channels][N][N] with x fastest and may be host or device
 * pointers (device data complete at the call, as for vk_gpe_set_psi).
 * Supported (width, modes): (32,16), (16,8), (8,4).
 *  vk_fno_layout: offsets of Wp, bp, {R, W, b} per layer, Q1, bq1, Q2, bq2
 *    and the total, in floats (returns the number written, 5 + 3*layers+2).
 *  vk_fno_set_params: FP32 values, rounded to BF16; resets Adam.
 *  vk_fno_get_params: the BF16 weights the network computes with.
 *  vk_fno_train_step: forward, mean-squared error against tgt [nb][2][N][N],
 *    backward (gradient readable by vk_fno_get_grad) and, for lr > 0, one
 *    Adam step (beta 0.9/0.999, eps 1e-8).
 *  vk_fno_predict_psi: one sample -> psi as double[2*N*N] (x, y), x fastest,
 *    ready for vk_gpe_set_psi; x = NULL uses the planes vk_fno_features built.
 *    flags VK_FNO_REAL keeps Re psi only (the ground state of a real
 *    potential is real; an imaginary error would excite phase modes that
 *    ARGLE damps only at rate alpha k^2); VK_FNO_ADD_TF adds input plane 2
 *    (the Thomas-Fermi profile) to Re psi, for a network trained on the
 *    correction psi - TF.
 *  vk_fno_features: input planes of one sample on the GPU: plane 0 =
 *    sum_v q_v exp(-r^2/(2 sigma^2)) over the vortices, plane 1 =
 *    (V0/vnorm) sum_p exp(-r^2/(2 d^2)) over the obstacles, with r from the
 *    particle potential's periodic coordinate (src/gpe_particle.f90:31-45);
 *    with cin >= 3, plane 2 = sqrt(max(0, (omega - V)/g)), V = vnorm plane 1;
 *    further planes are zero.  Kept in the handle and copied to x_out unless
 *    it is NULL.
 *  vk_fno_tf_guess: psi = sqrt(max(0, (omega - V)/g)) from plane 1. */
typedef struct vk_fno vk_fno;
int vk_fno_create(vk_fno** f, int N, int cin, int width, int modes, int layers, int proj, int batch);
void vk_fno_destroy(vk_fno* f);
long vk_fno_nparams(const vk_fno* f);
int vk_fno_layout(const vk_fno* f, long* offsets);
void* vk_fno_stream(const vk_fno* f);
int vk_fno_set_params(vk_fno* f, const float* params);
int vk_fno_get_params(vk_fno* f, float* params);
int vk_fno_get_grad(vk_fno* f, float* grad);
int vk_fno_forward(vk_fno* f, const float* x, float* y, int nb);
#define VK_FNO_REAL 1
#define VK_FNO_ADD_TF 2
int vk_fno_predict_psi(vk_fno* f, const float* x, double* psi_xy, int flags);
int vk_fno_features(vk_fno* f, int nv, const double* xv, const double* yv, const int* qv, double sigma,
                    int nobs, const double* pos, double V0, double d, double vnorm, double omega,
                    double g, float* x_out);
int vk_fno_tf_guess(vk_fno* f, double omega, double g, double vnorm, double* psi_xy);
int vk_fno_train_step(vk_fno* f, const float* x, const float* tgt, int nb, float lr, double* loss);

/* Spatio-temporal spectra (sts_gpe, src/diagnostics.f90:636-695, with the
 * post-processing of Run/CompSTS/driver.f90:134-265), resident on the
 * device: vk_sts_record snapshots the handle's kpsi (retained modes) after
 * the caller's steps; once nt snapshots are taken, vk_sts_compute applies
 * the Hann window (1 - cos(2 pi n/(nt-1)))/2, the forward DFT in time and
 * stsnorm |.|^2, stsnorm = (1/nt)^2 (1/N^2)^2, and returns
 *   spec2d[w][sh], sh = nint(|k|) in [0, vk_sts_shells()), summed per shell;
 *   spec1d[w][ky], the kx = 0 modes.
 * nt is a power of two in [32, 4096].  vk_sts_get_x0 returns the recorded
 * kpsi(ky, kx=0) columns (count x N complex), the content of the
 * reference's kpsi_x0 file.  An sts object uses its handle's stream and
 * kpsi: destroy it before the handle. */
typedef struct vk_sts vk_sts;
int vk_sts_create(vk_sts** s, vk_gpe* h, int nt);
void vk_sts_destroy(vk_sts* s);
int vk_sts_record(vk_sts* s);
int vk_sts_reset(vk_sts* s);
int vk_sts_count(const vk_sts* s);
int vk_sts_modes(const vk_sts* s);
int vk_sts_shells(const vk_sts* s);
int vk_sts_compute(vk_sts* s, double* spec2d, double* spec1d);
int vk_sts_get_x0(vk_sts* s, double* x0);
/* timing support: average duration (ms) of the compute kernels (time
 * transforms, shell sums, kx = 0 lines) over reps launches each */
int vk_sts_kernel_times(vk_sts* s, int reps, double* ms3);
/* snapshot n (0-based) as a full kpsi_t plane, zero off the retained modes:
 * the records of the reference's kpsi_nrst<n>p<r>.dat files. */
int vk_sts_get_kpsi(vk_sts* s, int n, double* kpsi_t);

/* Checkpoint files in the reference's formats (host only, no device).
 *  vk_wf_write: io_fields (src/io_fields.f90:41-50): nsplit files
 *    <dir>/wf<iopoint>p<r>.dat, r = 1..nsplit, Fortran unformatted
 *    sequential (4-byte record markers): header (INTEGER iopoint, REAL(8)
 *    time, REAL(8) dt), then N/nsplit records of N COMPLEX(8), one per row
 *    y of psi_xy (x fastest).
 *  vk_wf_read: restart_gpe / restart_from_argle_data
 *    (src/io_fields.f90:71-224, src/initial_config.f90:630-771): reads the
 *    nsplit_orig files of checkpoint nrestart into one field (regrouping).
 *  vk_part_checkpoint: gppart_checkpoint (src/gpe_particle.f90:144-184):
 *    appends 'iopoint ptime a b' lines to <dir>/gppos_ckp.dat_nrst<nrst>,
 *    gpvel_ckp..., gpf0ext_ckp...  (pos/vel/fext are (2,np) arrays).
 *  vk_part_restart: initialize_particles (src/gpe_particle.f90:92-141):
 *    reads np lines of <dir>/gppos.dat_nrst<nrst>, gpvel..., gpf0ext.... */
int vk_wf_write(const char* dir, int iopoint, double time, double dt, int N, int nsplit,
                const double* psi_xy);
int vk_wf_read(const char* dir, int nrestart, int N, int nsplit_orig, int* iopoint, double* time,
               double* dt, double* psi_xy);
int vk_part_checkpoint(const char* dir, int nrst, int iopoint, double ptime, int np, const double* pos,
                       const double* vel, const double* fext);
int vk_part_restart(const char* dir, int nrst, int np, double* pos, double* vel, double* fext);

const char* vk_last_error(void);
const char* vk_version(void);

#ifdef __cplusplus
}
#endif
#endif
