/*
 * CPU oracle for the 2D Gross-Pitaevskii pseudospectral step (test
 * infrastructure only: the tests compare the HIP library against it; it is
 * never the product path).
 *
 * Every function restates one piece of the reference Fortran, operation by
 * operation, and cites the file:line it follows (paths relative to the
 * reference repository root).
 *
 * Layouts are the reference's Fortran views, written as C arrays:
 *   x-space  psi(i1=x,i2=y)  -> psi[y*N + x]        (x fastest)
 *   k-space  kpsi(i2=ky,i1=kx) -> kpsi[kx*N + ky]   (ky fastest,
 *            FFTW_MPI_TRANSPOSED_OUT, src/mod_fft.f90:143-146,307-311)
 * Complex numbers are interleaved (re, im) doubles.
 */
#ifndef VK_ORACLE_H
#define VK_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex orc_c;

typedef struct {
    int N;            /* Nx = Ny */
    double L;         /* lengthx = two*pi */
    double dx, dvol;  /* src/mod_grid.f90:82-88 */
    double kalias;    /* REAL(N)/3 (src/mod_grid.f90:92) */
    int nshell;       /* int(1.414*Nx/2.0d0)+1 (src/mod_grid.f90:77) */
    double dt;
    double xi, alpha, g, omega, disp;   /* src/driver.f90:360-376 */
    /* particles (src/driver.f90:390-393, src/mod_gpparticles.f90) */
    int np;
    double V0, d, mass;
    int force_dealias;  /* 1: dealias rho-hat in the force with k1^2+k2^2 */
    /* ARGLE advection (src/gpe.f90:78-81,91-94,114-146): adv = 1 adds the
     * Taylor-Green term (GPE_ADVECTION); (vnx, vny) != 0 the counterflow
     * term (GPE_COUNTERFLOW, namelist gpcounterflow) */
    int adv;
    double vnx, vny;
    /* SGLE (GPE_SGLE, src/gpe.f90:149-217,226-266; namelist gpfiniteT,
     * src/driver.f90:59): noise of strength noisesigma =
     * sqrt((1/(2 alpha)) oneybeta) (src/driver.f90:376) and the chemical
     * potential relaxed towards norm = L^2 at rate nuN; ktrunc > 0 zeroes
     * kpsi beyond |k| = ktrunc (GPE_SELFTRUNC).  omega starts at chempot0
     * and is updated in place. */
    int sgle;
    double oneybeta, nuN, ktrunc;
} orc_params;

typedef struct orc_state orc_state;

/* --- grid / constants ------------------------------------------------ */
void orc_set_grid(int N, orc_params* p);
void orc_constants(double xi_in, double csound, double rho_avg,
                   double disp, orc_params* p);
void orc_kwn(int N, int* k);
int  orc_mask(const orc_params* p, int k1, int k2);
int  orc_shell(int k1, int k2);

/* --- FFT ---------------------------------------------------------------- */
void orc_twiddles(int N, orc_c* tw);
void orc_fft2d(int N, const orc_c* in, orc_c* out, int dir);
void orc_set_threads(int n);

/* --- state -------------------------------------------------------------- */
orc_state* orc_create(const orc_params* p);
void orc_destroy(orc_state* s);
orc_c* orc_psi(orc_state* s);
orc_c* orc_kpsi(orc_state* s);
orc_c* orc_vobj_k(orc_state* s);
orc_c* orc_trans_v(orc_state* s);
double* orc_part(orc_state* s, int which); /* 0 pos,1 vel,2 f0ext,3 frc */

void orc_sync_psi(orc_state* s);         /* kpsi=FFT(psi) */
void orc_ic_sync(orc_state* s);          /* dealias kpsi, psi=IFFT(kpsi) */
void orc_rk4_step(orc_state* s);
void orc_argle_step(orc_state* s, double* residual);
void orc_rhs_only(orc_state* s, orc_c* rhs_out);
void orc_diagnostics(orc_state* s, double* scal, double* spec, double* pscal);
void orc_particle_potential(orc_state* s);
void orc_translate_all(orc_state* s);
void orc_particle_force(orc_state* s, int ip, double* f);
void orc_shortrange(orc_state* s);
/* x-space noise field for the next SGLE step: complex standard normals
 * (v1 + i v2 per point, before the noisesigma/sqrt(dvol) scale) */
void orc_set_noise(orc_state* s, const orc_c* noise);
/* caller-supplied advection velocity (adv = 2): u = (ux, uy) per point,
 * x-space, in place of the Taylor-Green field */
void orc_set_velocity(orc_state* s, const double* ux, const double* uy);
double orc_omega(const orc_state* s);

#ifdef __cplusplus
}
#endif
#endif
