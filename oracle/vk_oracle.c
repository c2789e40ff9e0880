/*
 * CPU oracle: a plain-C restatement of the reference's 2D GPE path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * reference the tests compare against (or the timed CPU baseline) -- never
 * as the product path.
 *
 * Each function cites the reference file:line it follows.  The reference
 * uses FFTW3-MPI for its transforms (an external library, not vendored,
 * version 3.3.x per src/Makefile:97-136); here the transform is this file's
 * own radix-2 FFT with the same conventions: unnormalised, sign -1 forward,
 * the inverse scaled by fftwscale = 1/(Nx*Ny) (src/mod_fft.f90:414,486).
 *
 * Pinning (what checks each function): see the header of tests/test_oracle.py
 * and the oracle section of the design notes.  In short: the FFT is pinned by
 * a brute-force DFT and numpy.fft fixtures; masks/shells by integer
 * identities; RK4/ARGLE by exact per-mode closed forms (g=0) and plane-wave
 * solutions; spectra by Helmholtz purity, single-mode placement and
 * Parseval; particles by closed forms, grid shifts and a brute-force
 * real-space force sum.
 */
#include "vk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RE(z) creal(z)
#define IM(z) cimag(z)

static inline orc_c mkc(double a, double b) { return CMPLX(a, b); }
/* complex product written out: (a+bi)(c+di) = (ac-bd) + (ad+bc)i */
static inline orc_c cmul(orc_c x, orc_c y)
{
    double a = RE(x), b = IM(x), c = RE(y), d = IM(y);
    return mkc(a * c - b * d, a * d + b * c);
}
static inline orc_c rmul(double r, orc_c x) { return mkc(r * RE(x), r * IM(x)); }
static inline orc_c cadd(orc_c x, orc_c y) { return mkc(RE(x) + RE(y), IM(x) + IM(y)); }
static inline orc_c csub(orc_c x, orc_c y) { return mkc(RE(x) - RE(y), IM(x) - IM(y)); }
/* Fortran ABS(complex) is hypot(re, im) */
static inline double cabs_f(orc_c x) { return hypot(RE(x), IM(x)); }
static inline double sq(double x) { return x * x; }

/* ----------------------------------------------------------------------- */
/* Grid and constants                                                       */
/* ----------------------------------------------------------------------- */

/* src/mod_grid.f90:63-101 (set_grid).  1.414 is a default-real literal, so
 * 1.414*Nx is evaluated in single precision before /2.0d0. */
void orc_set_grid(int N, orc_params* p)
{
    const double pi = 4.0 * atan(1.0);      /* src/mod_constants.f90:16 */
    p->N = N;
    float a = 1.414f * (float)N;
    p->nshell = (int)((double)a / 2.0) + 1;
    p->L = 2.0 * pi;
    p->dx = p->L / N;
    p->dvol = p->dx * p->dx;
    p->kalias = (double)N / 3.0;
}

/* src/driver.f90:360-376: xi <- xi*dx; alpha = csound*xi*SQRT(rho_avg/two);
 * gstrength = csound**2/(two*alpha); omega_chem = chempot0 = gstrength. */
void orc_constants(double xi_in, double csound, double rho_avg, double disp,
                   orc_params* p)
{
    p->xi = xi_in * p->dx;
    p->alpha = csound * p->xi * sqrt(rho_avg / 2.0);
    p->g = (csound * csound) / (2.0 * p->alpha);
    p->omega = p->g;
    p->disp = disp;
}

/* src/pspectral.f90:22,27: kwn(i) = (i-1) - N*(i/(N/2+1)), i = 1..N */
void orc_kwn(int N, int* k)
{
    for (int i = 1; i <= N; ++i) k[i - 1] = (i - 1) - N * (i / (N / 2 + 1));
}

/* Dealias rule, e.g. src/gpe.f90:346: zero iff ksqr .GE. kalias**2, with
 * ksqr an INTEGER and kalias a REAL(GP); evaluated exactly as written. */
int orc_mask(const orc_params* p, int k1, int k2)
{
    int ksqr = k1 * k1 + k2 * k2;
    return (double)ksqr >= p->kalias * p->kalias;
}

/* src/diagnostics.f90:341-343: rk = sqrt(kx**2+ky**2); mshl = nint(rk). */
int orc_shell(int k1, int k2)
{
    double kx = (double)k1, ky = (double)k2;
    double rk = sqrt(kx * kx + ky * ky);
    return (int)lround(rk);
}

/* ----------------------------------------------------------------------- */
/* FFT                                                                      */
/* ----------------------------------------------------------------------- */

static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

/* w_k = exp(-2 pi i k / N), k = 0..N-1, from long-double sin/cos of the
 * argument reduced to the first octant (exact symmetries elsewhere). */
void orc_twiddles(int N, orc_c* tw)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < N; ++k) {
        long double c, s;
        if (N < 4) {
            long double t = 2.0L * pi * (long double)k / (long double)N;
            c = cosl(t);
            s = sinl(t);
        } else {
            /* angle = quad*pi/2 + 2 pi a/N with a in [0, N/4) */
            int q = N / 4, quad = k / q, a = k % q, flip = 0;
            if (2 * a > q) { a = q - a; flip = 1; }   /* keep t <= pi/4 */
            long double t = 2.0L * pi * (long double)a / (long double)N;
            c = cosl(t);
            s = sinl(t);
            if (flip) { long double tmp = c; c = s; s = tmp; }
            for (int i = 0; i < quad; ++i) { long double tmp = c; c = -s; s = tmp; }
        }
        tw[k] = mkc((double)c, -(double)s);
    }
}

/* In-place iterative radix-2 FFT of length N (power of two), unnormalised,
 * sign -1 (sign=+1 uses conj twiddles). */
static void fft1d(orc_c* a, int N, const orc_c* tw, int sign)
{
    for (int i = 1, j = 0; i < N; ++i) {
        int bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { orc_c t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= N; len <<= 1) {
        int half = len >> 1, step = N / len;
        for (int i = 0; i < N; i += len)
            for (int j = 0; j < half; ++j) {
                orc_c w = tw[j * step];
                if (sign > 0) w = conj(w);
                orc_c u = a[i + j], v = cmul(a[i + j + half], w);
                a[i + j] = cadd(u, v);
                a[i + j + half] = csub(u, v);
            }
    }
}

/* 2D transform with the reference's plan conventions
 * (src/mod_fft.f90:307-311 plans, :482-487 execution + scaling):
 *  dir = -1: in = x-space [y][x] -> out = k-space [kx][ky], unnormalised
 *  dir = +1: in = k-space [kx][ky] -> out = x-space [y][x], times 1/(N*N)
 * in == out is allowed (the reference's gpnlin/cmpxtmp/tmp buffers alias
 * their physical and Fourier views). */
void orc_fft2d(int N, const orc_c* in, orc_c* out, int dir)
{
    size_t n2 = (size_t)N * N;
    orc_c* tw = (orc_c*)malloc(sizeof(orc_c) * N);
    orc_c* tmp = (orc_c*)malloc(sizeof(orc_c) * n2);
    orc_twiddles(N, tw);
    /* pass 1: transform along the contiguous axis of the input, B rows at a
     * time, transposed on store (tmp[c][r]) in runs of B contiguous values */
    const int B = N < 16 ? N : 16;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int r0 = 0; r0 < N; r0 += B) {
        orc_c* rows = (orc_c*)malloc(sizeof(orc_c) * (size_t)B * N);
        for (int b = 0; b < B; ++b) {
            memcpy(rows + (size_t)b * N, in + (size_t)(r0 + b) * N, sizeof(orc_c) * N);
            fft1d(rows + (size_t)b * N, N, tw, dir);
        }
        for (int c = 0; c < N; ++c)
            for (int b = 0; b < B; ++b) tmp[(size_t)c * N + r0 + b] = rows[(size_t)b * N + c];
        free(rows);
    }
    /* pass 2: transform along the other axis (now contiguous), stored in the
     * transposed orientation of the input, which is exactly the output
     * layout in both directions. */
    const double scale = 1.0 / ((double)N * (double)N);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int c = 0; c < N; ++c) {
        orc_c* row = tmp + (size_t)c * N;
        fft1d(row, N, tw, dir);
    }
    /* tmp[c][r] now holds: dir=-1, input [y][x]: pass 1 along x gives row y,
     * col kx -> tmp[kx][y]; pass 2 along y -> tmp[kx][ky] = out layout.
     * dir=+1, input [kx][ky]: pass 1 along ky -> tmp[y][kx]; pass 2 along
     * kx -> tmp[y][x] = out layout. */
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (size_t i = 0; i < n2; ++i)
        out[i] = (dir > 0) ? rmul(scale, tmp[i]) : tmp[i];
    free(tmp);
    free(tw);
}

/* ----------------------------------------------------------------------- */
/* State                                                                    */
/* ----------------------------------------------------------------------- */

struct orc_state {
    orc_params p;
    int* kwn;
    orc_c *psi, *kpsi, *nlin, *str1, *str2, *tmp1, *tmp2, *tmp3, *cmpxtmp;
    orc_c *vobj_k, *trans_v;     /* cgp_four_vobj, cgp_phys_trans_vobj */
    /* particles: (idmn, ip) -> [ip*2 + idmn] */
    double *pos, *vel, *f0ext, *frc, *tpos, *tvel, *kq[4], *ku[4], *sr;
    int sr_init;                 /* shortrange_repulsion has run */
    double gpdeltaE, gpr0obj;
    orc_c* noise;                /* SGLE noise of the next step (x-space) */
    double *ux, *uy;             /* advection velocity for adv = 2 */
};

static orc_c* calloc_c(size_t n) { return (orc_c*)calloc(n, sizeof(orc_c)); }

orc_state* orc_create(const orc_params* p)
{
    orc_state* s = (orc_state*)calloc(1, sizeof(orc_state));
    size_t n2 = (size_t)p->N * p->N;
    s->p = *p;
    s->kwn = (int*)malloc(sizeof(int) * p->N);
    orc_kwn(p->N, s->kwn);
    s->psi = calloc_c(n2); s->kpsi = calloc_c(n2); s->nlin = calloc_c(n2);
    s->str1 = calloc_c(n2); s->str2 = calloc_c(n2); s->tmp1 = calloc_c(n2);
    s->tmp2 = calloc_c(n2); s->tmp3 = calloc_c(n2); s->cmpxtmp = calloc_c(n2);
    s->vobj_k = calloc_c(n2); s->trans_v = calloc_c(n2);
    s->noise = calloc_c(n2);
    s->ux = (double*)calloc(n2, sizeof(double)); s->uy = (double*)calloc(n2, sizeof(double));
    int np = p->np > 0 ? p->np : 1;
    s->pos = calloc(2 * np, sizeof(double)); s->vel = calloc(2 * np, sizeof(double));
    s->f0ext = calloc(2 * np, sizeof(double)); s->frc = calloc(2 * np, sizeof(double));
    s->tpos = calloc(2 * np, sizeof(double)); s->tvel = calloc(2 * np, sizeof(double));
    s->sr = calloc(2 * np, sizeof(double));
    for (int i = 0; i < 4; ++i) {
        s->kq[i] = calloc(2 * np, sizeof(double));
        s->ku[i] = calloc(2 * np, sizeof(double));
    }
    return s;
}

void orc_destroy(orc_state* s)
{
    if (!s) return;
    free(s->kwn);
    free(s->psi); free(s->kpsi); free(s->nlin); free(s->str1); free(s->str2);
    free(s->tmp1); free(s->tmp2); free(s->tmp3); free(s->cmpxtmp);
    free(s->vobj_k); free(s->trans_v);
    free(s->noise); free(s->ux); free(s->uy);
    free(s->pos); free(s->vel); free(s->f0ext); free(s->frc);
    free(s->tpos); free(s->tvel); free(s->sr);
    for (int i = 0; i < 4; ++i) { free(s->kq[i]); free(s->ku[i]); }
    free(s);
}

void orc_set_noise(orc_state* s, const orc_c* noise)
{
    memcpy(s->noise, noise, sizeof(orc_c) * (size_t)s->p.N * s->p.N);
}
void orc_set_velocity(orc_state* s, const double* ux, const double* uy)
{
    memcpy(s->ux, ux, sizeof(double) * (size_t)s->p.N * s->p.N);
    memcpy(s->uy, uy, sizeof(double) * (size_t)s->p.N * s->p.N);
}
double orc_omega(const orc_state* s) { return s->p.omega; }
orc_c* orc_psi(orc_state* s) { return s->psi; }
orc_c* orc_kpsi(orc_state* s) { return s->kpsi; }
orc_c* orc_vobj_k(orc_state* s) { return s->vobj_k; }
orc_c* orc_trans_v(orc_state* s) { return s->trans_v; }
double* orc_part(orc_state* s, int which)
{
    switch (which) {
    case 0: return s->pos;
    case 1: return s->vel;
    case 2: return s->f0ext;
    case 3: return s->frc;
    default: return NULL;
    }
}

#define FOR_K(s)                                                            \
    for (int i1 = 0; i1 < (s)->p.N; ++i1)                                   \
        for (int i2 = 0; i2 < (s)->p.N; ++i2)
/* the same loop over independent modes (no reduction), split over threads */
#define FOR_K_PAR(s)                                                        \
    _Pragma("omp parallel for num_threads(g_threads) schedule(static)")      \
    for (int i1 = 0; i1 < (s)->p.N; ++i1)                                   \
        for (int i2 = 0; i2 < (s)->p.N; ++i2)
#define FOR_X_PAR(n2)                                                       \
    _Pragma("omp parallel for num_threads(g_threads) schedule(static)")      \
    for (size_t i = 0; i < (n2); ++i)
#define KIDX(s) ((size_t)i1 * (s)->p.N + i2)
#define K1 (s->kwn[i1])
#define K2 (s->kwn[i2])

/* restart_from_argle_data: psi -> kpsi (src/initial_config.f90:759) */
void orc_sync_psi(orc_state* s)
{
    orc_fft2d(s->p.N, s->psi, s->kpsi, -1);
}

/* ic_otoc_gpe with k_otoc never matching (src/initial_config.f90:316-343):
 * dealias kpsi, then psi = IFFT(kpsi). */
void orc_ic_sync(orc_state* s)
{
    FOR_K_PAR(s) {
        if (orc_mask(&s->p, K1, K2)) s->kpsi[KIDX(s)] = 0;
    }
    orc_fft2d(s->p.N, s->kpsi, s->psi, +1);
}

/* ----------------------------------------------------------------------- */
/* Particles                                                                */
/* ----------------------------------------------------------------------- */

/* src/gpe_particle.f90:3-90 (particle_potential) */
void orc_particle_potential(orc_state* s)
{
    const orc_params* p = &s->p;
    const int N = p->N;
    const double pi = 4.0 * atan(1.0);
    double c1 = 2 * pi / (2.0 * p->L);
    double c2 = 2 * pi / (2.0 * p->L);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i2 = 0; i2 < N; ++i2) {
        double y = i2 * p->dx;
        for (int i1 = 0; i1 < N; ++i1) {
            double x = i1 * p->dx;
            double sx = sin(c1 * x), sy = sin(c2 * y);
            double sx2 = sx * sx, sy2 = sy * sy;
            double xprd = (1.0 / c1) * (sx + (1.0 / 6.0) * (sx2 * sx) + (1.0 / 40.0) * (sx2 * sx2 * sx));
            double yprd = (1.0 / c2) * (sy + (1.0 / 6.0) * (sy2 * sy) + (1.0 / 40.0) * (sy2 * sy2 * sy));
            double rprd = sqrt(xprd * xprd + yprd * yprd);
            double v = p->V0 * exp(-(rprd * rprd) / (2.0 * (p->d * p->d)));
            s->cmpxtmp[(size_t)i2 * N + i1] = mkc(v, 0.0);
        }
    }
    orc_fft2d(N, s->cmpxtmp, s->cmpxtmp, -1);
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        if (orc_mask(p, K1, K2)) s->cmpxtmp[k] = 0;
        s->vobj_k[k] = s->cmpxtmp[k];
        s->cmpxtmp[k] = 0;
    }
}

/* src/gpe_particle.f90:327-374 (translate_particle_potential(0)):
 * V = IFFT(sum_p exp(-i(kx qx + ky qy)) Vhat) into cgp_phys_trans_vobj. */
void orc_translate_all(orc_state* s)
{
    const orc_params* p = &s->p;
    FOR_K_PAR(s) {
        double kx = K1, ky = K2;
        double sr = 0.0, si = 0.0;
        for (int ip = 0; ip < p->np; ++ip) {
            double th = kx * s->pos[2 * ip] + ky * s->pos[2 * ip + 1];
            sr += cos(th);
            si += -sin(th);
        }
        s->trans_v[KIDX(s)] = cmul(mkc(sr, si), s->vobj_k[KIDX(s)]);
    }
    orc_fft2d(p->N, s->trans_v, s->trans_v, +1);
}

static double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

/* src/gpe_particle.f90:378-427 (shortrange_repulsion).  gpdeltaE = 0.062 is
 * a default-real literal. */
void orc_shortrange(orc_state* s)
{
    const orc_params* p = &s->p;
    const double pi = 4.0 * atan(1.0);
    s->gpdeltaE = (double)0.062f;
    s->gpr0obj = 4.0 * p->xi;
    s->sr_init = 1;
    double r0obj12 = ipow(s->gpr0obj, 12);
    double rf = 384.0 * s->gpdeltaE * ipow(pi, 12) * r0obj12 / ipow(p->L, 12);
    for (int i = 0; i < 2 * p->np; ++i) s->sr[i] = 0.0;
    for (int a = 0; a < p->np; ++a)
        for (int b = 0; b < p->np; ++b) {
            double x = s->pos[2 * a] - s->pos[2 * b];
            double y = s->pos[2 * a + 1] - s->pos[2 * b + 1];
            double den = ipow(-3.0 + cos(x) + cos(y), 7);
            double fx = -rf * sin(x) / den;
            double fy = -rf * sin(y) / den;
            if (a == b) { fx = 0.0; fy = 0.0; }
            s->sr[2 * a] += fx;
            s->sr[2 * a + 1] += fy;
        }
    if (p->np == 1) { s->sr[0] = 0.0; s->sr[1] = 0.0; }
}

/* src/gpe_particle.f90:219-324 (translate_particle_potential(ip), ip>=1):
 * force on particle ip (0-based here).  Note the reference leaves ksqr
 * uninitialised at :275; force_dealias selects the evident intent (mask by
 * k1^2+k2^2) or no dealiasing. */
void orc_particle_force(orc_state* s, int ip, double* f)
{
    const orc_params* p = &s->p;
    const int N = p->N;
    size_t n2 = (size_t)N * N;
    FOR_K_PAR(s) {
        double th = (double)K1 * s->pos[2 * ip] + (double)K2 * s->pos[2 * ip + 1];
        s->trans_v[KIDX(s)] = cmul(mkc(cos(th), -sin(th)), s->vobj_k[KIDX(s)]);
    }
    orc_fft2d(N, s->trans_v, s->trans_v, +1);
    orc_shortrange(s);
    FOR_X_PAR(n2) s->cmpxtmp[i] = mkc(sq(cabs_f(s->psi[i])), 0.0);
    orc_fft2d(N, s->cmpxtmp, s->cmpxtmp, -1);
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        if (p->force_dealias && orc_mask(p, K1, K2)) s->cmpxtmp[k] = 0;
        orc_c c = s->cmpxtmp[k];
        s->tmp1[k] = cmul(mkc(0.0, (double)K1), c);
        s->tmp2[k] = cmul(mkc(0.0, (double)K2), c);
    }
    orc_fft2d(N, s->tmp1, s->tmp1, +1);
    orc_fft2d(N, s->tmp2, s->tmp2, +1);
    double fx = 0.0, fy = 0.0;
    for (size_t i = 0; i < n2; ++i) {
        fx = fx - RE(cmul(rmul(2.0 * p->alpha, s->tmp1[i]), s->trans_v[i]));
        fy = fy - RE(cmul(rmul(2.0 * p->alpha, s->tmp2[i]), s->trans_v[i]));
    }
    fx *= p->dvol;
    fy *= p->dvol;
    f[0] = fx + s->sr[2 * ip] + s->f0ext[2 * ip];
    f[1] = fy + s->sr[2 * ip + 1] + s->f0ext[2 * ip + 1];
}

/* src/gpe_particle.f90:429-531 (gp_particle_motion_rk4).  Particles are
 * advanced one at a time, so the force on particle ip sees the already
 * updated positions of particles < ip through shortrange_repulsion. */
static void particle_rk4_stage(orc_state* s, int stage)
{
    const orc_params* p = &s->p;
    double f[2];
    if (stage == 1)
        for (int i = 0; i < 2 * p->np; ++i) { s->tvel[i] = s->vel[i]; s->tpos[i] = s->pos[i]; }
    for (int ip = 0; ip < p->np; ++ip) {
        orc_particle_force(s, ip, f);
        for (int d = 0; d < 2; ++d) {
            int i = 2 * ip + d;
            double ku = p->dt * f[d] / p->mass;
            double kq = p->dt * s->vel[i];
            s->ku[stage - 1][i] = ku;
            s->kq[stage - 1][i] = kq;
            if (stage == 1 || stage == 2) {
                s->vel[i] = s->tvel[i] + 0.5 * ku;
                s->pos[i] = s->tpos[i] + 0.5 * kq;
            } else if (stage == 3) {
                s->vel[i] = s->tvel[i] + ku;
                s->pos[i] = s->tpos[i] + kq;
            } else {
                s->vel[i] = s->tvel[i] + (1.0 / 6.0) * (s->ku[0][i] + 2.0 * s->ku[1][i] + 2.0 * s->ku[2][i] + s->ku[3][i]);
                s->pos[i] = s->tpos[i] + (1.0 / 6.0) * (s->kq[0][i] + 2.0 * s->kq[1][i] + 2.0 * s->kq[2][i] + s->kq[3][i]);
            }
        }
    }
}

/* ----------------------------------------------------------------------- */
/* RHS, RK4, ARGLE                                                          */
/* ----------------------------------------------------------------------- */

/* src/gpe.f90:498-671 (gpe_rhs, without GPE_FRCDISP).  Result in nlin
 * (cgp_four_nlin); tmp1 holds (alpha*k^2-omega)*kpsi. */
static void gpe_rhs(orc_state* s)
{
    const orc_params* p = &s->p;
    const int N = p->N;
    size_t n2 = (size_t)N * N;
    /* prefacrhs = (one - zi*disp_factor)/zi  (src/gpe.f90:529) */
    orc_c num = mkc(1.0, -p->disp);
    orc_c pref = mkc(IM(num), -RE(num));  /* (a+bi)/i = b - ai */
    if (p->np > 0) orc_translate_all(s);
    FOR_X_PAR(n2) s->nlin[i] = mkc(sq(cabs_f(s->psi[i])), 0.0);
    orc_fft2d(N, s->nlin, s->nlin, -1);
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        double kx = K1, ky = K2;
        double kr2 = kx * kx + ky * ky;
        s->tmp1[k] = rmul(p->alpha * kr2 - p->omega, s->kpsi[k]);
        if (orc_mask(p, K1, K2)) s->nlin[k] = 0;
    }
    orc_fft2d(N, s->nlin, s->nlin, +1);
    FOR_X_PAR(n2) {
        orc_c v = cmul(rmul(p->g, s->nlin[i]), s->psi[i]);
        if (p->np > 0) v = cadd(v, cmul(s->trans_v[i], s->psi[i]));
        s->nlin[i] = v;
    }
    orc_fft2d(N, s->nlin, s->nlin, -1);
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        s->nlin[k] = cmul(pref, cadd(s->tmp1[k], s->nlin[k]));
        if (orc_mask(p, K1, K2)) s->nlin[k] = 0;
    }
}

void orc_rhs_only(orc_state* s, orc_c* rhs_out)
{
    gpe_rhs(s);
    memcpy(rhs_out, s->nlin, sizeof(orc_c) * (size_t)s->p.N * s->p.N);
}

/* src/gpe.f90:310-494 (gpe_rk4) */
void orc_rk4_step(orc_state* s)
{
    const orc_params* p = &s->p;
    const int N = p->N;
    const double dt = p->dt;
    const double oney6 = 1.0 / 6.0, half = 0.5, two = 2.0;
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        if (orc_mask(p, K1, K2)) s->kpsi[k] = 0;
        s->str1[k] = s->kpsi[k];
        s->str2[k] = 0;
    }
    for (int stage = 1; stage <= 4; ++stage) {
        gpe_rhs(s);
        if (p->np > 0) particle_rk4_stage(s, stage);
        double wacc = stage == 1 || stage == 4 ? dt * oney6 : dt * oney6 * two;
        FOR_K_PAR(s) {
            size_t k = KIDX(s);
            orc_c r = s->nlin[k];
            s->str2[k] = cadd(s->str2[k], rmul(wacc, r));
            if (stage == 1 || stage == 2)
                s->kpsi[k] = cadd(s->str1[k], rmul(dt * half, r));
            else if (stage == 3)
                s->kpsi[k] = cadd(s->str1[k], rmul(dt, r));
            else
                s->kpsi[k] = cadd(s->str1[k], s->str2[k]);
            if (orc_mask(p, K1, K2)) s->kpsi[k] = 0;
        }
        orc_fft2d(N, s->kpsi, s->psi, +1);
    }
}

/* src/gpe.f90:5-224 (ARGLE; with the advective terms of GPE_ADVECTION and
 * GPE_COUNTERFLOW when p->adv / p->vnx, p->vny are set, adv = 2 taking the
 * velocity from orc_set_velocity in place of the Taylor-Green field; with
 * the GPE_SGLE noise and chemical-potential update when p->sgle is set).
 * residual (not defined by the reference, which has no convergence test):
 *   ||kpsi_new - kpsi_old||_2 / (dt * ||kpsi_old||_2) over all modes. */
void orc_argle_step(orc_state* s, double* residual)
{
    const orc_params* p = &s->p;
    const int N = p->N;
    size_t n2 = (size_t)N * N;
    const double dt = p->dt;
    const int cflow = p->vnx != 0.0 || p->vny != 0.0;
    orc_fft2d(N, s->kpsi, s->psi, +1);
    if (p->adv || cflow) {
        /* gradients of psi: IFFT(i kx kpsi), IFFT(i ky kpsi) (:78-81,91-94) */
        FOR_K_PAR(s) {
            size_t k = KIDX(s);
            s->tmp1[k] = cmul(mkc(0.0, (double)K1), s->kpsi[k]);
            s->tmp2[k] = cmul(mkc(0.0, (double)K2), s->kpsi[k]);
        }
        orc_fft2d(N, s->tmp1, s->tmp1, +1);
        orc_fft2d(N, s->tmp2, s->tmp2, +1);
    }
    if (p->np > 0) orc_translate_all(s);
    FOR_X_PAR(n2) s->nlin[i] = mkc(sq(cabs_f(s->psi[i])), 0.0);
    orc_fft2d(N, s->nlin, s->nlin, -1);
    FOR_K_PAR(s) {
        if (orc_mask(p, K1, K2)) s->nlin[KIDX(s)] = 0;
    }
    orc_fft2d(N, s->nlin, s->nlin, +1);
    FOR_X_PAR(n2) {
        orc_c v = cmul(rmul(-p->g, s->nlin[i]), s->psi[i]);
        if (p->np > 0) {
            orc_c pot = cmul(s->trans_v[i], s->psi[i]);
            v = cadd(v, mkc(-RE(pot), -IM(pot)));
        }
        if (p->adv) {
            /* Taylor-Green velocity (:114-130), or the supplied field:
             * -i (u.grad psi) - (1/4)(1/alpha)|u|^2 psi */
            const double x = (double)(i % N) * p->dx, y = (double)(i / N) * p->dx;
            const double ux = p->adv == 2 ? s->ux[i] : sin(x) * cos(y);
            const double uy = p->adv == 2 ? s->uy[i] : -cos(x) * sin(y);
            const orc_c a = cadd(rmul(ux, s->tmp1[i]), rmul(uy, s->tmp2[i]));
            v = csub(v, cmul(mkc(0.0, 1.0), a));
            v = csub(v, rmul(0.25 * (1.0 / p->alpha) * (ux * ux + uy * uy), s->psi[i]));
        }
        if (cflow) {
            /* constant counterflow (:136-146): -i (v.grad psi) */
            const orc_c a = cadd(rmul(p->vnx, s->tmp1[i]), rmul(p->vny, s->tmp2[i]));
            v = csub(v, cmul(mkc(0.0, 1.0), a));
        }
        s->nlin[i] = v;
    }
    orc_fft2d(N, s->nlin, s->nlin, -1);
    if (p->sgle) {
        /* noise (:150-164): CMPLX(v1,v2)*(1/SQRT(dvol))*noisesigma, from the
         * supplied standard normals */
        const double noisesigma = sqrt((1.0 / (2.0 * p->alpha)) * p->oneybeta);
        FOR_X_PAR(n2) s->cmpxtmp[i] = rmul(noisesigma, rmul(1.0 / sqrt(p->dvol), s->noise[i]));
        /* finiteT_adhoc_chempot (:226-266) on psi = IFFT(kpsi) of this step:
         * norm = sum |psi|^2 dvol; chempot -= dt (nuN/(Lx Ly)) (norm - Nparticle),
         * Nparticle = Lx Ly (src/driver.f90:362) */
        double norm = 0.0;
        for (size_t i = 0; i < n2; ++i) norm = norm + sq(cabs_f(s->psi[i]));
        norm = norm * p->dvol;
        s->p.omega = p->omega - dt * (p->nuN / (p->L * p->L)) * (norm - p->L * p->L);
        orc_fft2d(N, s->cmpxtmp, s->cmpxtmp, -1);
    }
    /* the update per mode (into tmp3), then the residual sums in the loop
     * order of the update, then kpsi <- tmp3.  SGLE (:200-217): the noise
     * term FFT(noise) sqrt(dt), dealiased, is added to the numerator; with
     * ktrunc > 0 modes with ksqr > ktrunc^2 are zeroed (GPE_SELFTRUNC). */
    const double sdt = sqrt(dt);
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        double kx = K1, ky = K2;
        double kr2 = kx * kx + ky * ky;
        orc_c old = s->kpsi[k];
        orc_c t1 = cadd(s->nlin[k], rmul(p->omega, old));
        double t2 = 1.0 + p->alpha * kr2 * dt;
        orc_c num = cadd(old, rmul(dt, t1));
        if (p->sgle) {
            orc_c nt = rmul(sdt, s->cmpxtmp[k]);
            if (orc_mask(p, K1, K2)) nt = 0;
            num = cadd(num, nt);
        }
        orc_c nw = mkc(RE(num) / t2, IM(num) / t2);
        if (p->sgle && p->ktrunc > 0.0 && (double)(K1 * K1 + K2 * K2) > p->ktrunc * p->ktrunc) nw = 0;
        s->tmp3[k] = nw;
    }
    double dnum = 0.0, dden = 0.0;
    for (size_t k = 0; k < n2; ++k) {
        orc_c old = s->kpsi[k], nw = s->tmp3[k];
        dnum += sq(RE(nw) - RE(old)) + sq(IM(nw) - IM(old));
        dden += sq(RE(old)) + sq(IM(old));
    }
    memcpy(s->kpsi, s->tmp3, sizeof(orc_c) * n2);
    if (residual) *residual = sqrt(dnum) / (dt * sqrt(dden));
}

/* ----------------------------------------------------------------------- */
/* Diagnostics                                                              */
/* ----------------------------------------------------------------------- */

/* src/diagnostics.f90:54-426 (diagnostics_gpe, energies + spectra).
 * scal[8] = {egrad, echem, eintr, etot, norm, momx, momy, n0 = spec(0,6)}.
 * spec is spec_gp(0:nshell, 1:6) in Fortran order: spec[(i-1)*(nshell+1)+m].
 * pscal[5] (particles) = {eppot, eobj_kin, esrep, momobj_x, momobj_y}; the
 * per-particle forces (gp_frc_obj, :528-533) are stored in orc_part(s,3). */
void orc_diagnostics(orc_state* s, double* scal, double* spec, double* pscal)
{
    const orc_params* p = &s->p;
    const int N = p->N;
    size_t n2 = (size_t)N * N;
    const double two = 2.0, half = 0.5, alpha = p->alpha;
    const double smlnum = (double)1e-30f;   /* :43, default-real literal */
    const double L2 = p->L * p->L;
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        s->tmp1[k] = cmul(mkc(0.0, (double)K1), s->kpsi[k]);
        s->tmp2[k] = cmul(mkc(0.0, (double)K2), s->kpsi[k]);
    }
    orc_fft2d(N, s->tmp1, s->tmp1, +1);
    orc_fft2d(N, s->tmp2, s->tmp2, +1);
    if (p->np > 0) orc_translate_all(s);
    double momx = 0.0, momy = 0.0;
    for (size_t i = 0; i < n2; ++i) {
        orc_c cp = conj(s->psi[i]);
        momx += IM(cmul(cp, s->tmp1[i]));
        momy += IM(cmul(cp, s->tmp2[i]));
        s->nlin[i] = mkc(sq(cabs_f(s->psi[i])), 0.0);
    }
    momx = two * alpha * momx * p->dvol / L2;
    momy = two * alpha * momy * p->dvol / L2;
    orc_fft2d(N, s->nlin, s->nlin, -1);
    FOR_K_PAR(s) {
        size_t k = KIDX(s);
        if (orc_mask(p, K1, K2)) s->nlin[k] = 0;
        s->str1[k] = s->nlin[k];
    }
    orc_fft2d(N, s->nlin, s->nlin, +1);
    double egrad = 0, echem = 0, eintr = 0, norm = 0, eppot = 0;
    for (size_t i = 0; i < n2; ++i) {
        double ap2 = sq(cabs_f(s->psi[i]));
        egrad = egrad + alpha * (sq(cabs_f(s->tmp1[i])) + sq(cabs_f(s->tmp2[i])));
        echem = echem + p->omega * ap2;
        eintr = eintr + half * p->g * cabs_f(s->nlin[i]) * ap2;
        if (p->np > 0) eppot = eppot + RE(s->trans_v[i]) * RE(s->nlin[i]);
        norm = norm + ap2;
    }
    egrad = two * alpha * egrad * p->dvol / L2;
    echem = two * alpha * echem * p->dvol / L2;
    eintr = two * alpha * eintr * p->dvol / L2;
    eppot = two * alpha * eppot * p->dvol / L2;
    norm = norm * p->dvol;
    /* particle energies (:210-239); gpdeltaE/gpr0obj are only set once
     * shortrange_repulsion has run, zero before that. */
    if (p->np > 0 && pscal) {
        const double pi = 4.0 * atan(1.0);
        double fac = 64.0 * s->gpdeltaE * ipow(pi, 12) * ipow(s->gpr0obj, 12) / ipow(p->L, 12);
        double ekin = 0.0, esrep = 0.0, mo[2] = {0.0, 0.0};
        for (int a = 0; a < p->np; ++a) {
            for (int d = 0; d < 2; ++d) {
                ekin = ekin + sq(s->vel[2 * a + d]);
                mo[d] = mo[d] + p->mass * s->vel[2 * a + d];
            }
            for (int b = 0; b < p->np; ++b) {
                double x = s->pos[2 * a] - s->pos[2 * b];
                double y = s->pos[2 * a + 1] - s->pos[2 * b + 1];
                double erep = fac / ipow(-3.0 + cos(x) + cos(y), 6);
                if (a == b) erep = 0.0;
                esrep = esrep + erep;
            }
        }
        ekin = half * p->mass * ekin;
        esrep = half * esrep;
        if (p->np == 1) esrep = 0.0;
        pscal[0] = eppot; pscal[1] = ekin; pscal[2] = esrep;
        pscal[3] = mo[0]; pscal[4] = mo[1];
    }
    /* :265-317  w = 2 alpha Im(psi* grad psi)/(|psi|+smlnum),
     *           dabs = Re(psi* grad psi)/(|psi|+smlnum) */
    FOR_X_PAR(n2) {
        double rholoc = cabs_f(s->psi[i]) + smlnum;
        orc_c cp = conj(s->psi[i]);
        orc_c a1 = cmul(cp, s->tmp1[i]);
        orc_c a2 = cmul(cp, s->tmp2[i]);
        a1 = mkc(RE(a1) / rholoc, IM(a1) / rholoc);
        a2 = mkc(RE(a2) / rholoc, IM(a2) / rholoc);
        s->cmpxtmp[i] = mkc(RE(a1), 0.0);
        s->tmp3[i] = mkc(RE(a2), 0.0);
        s->tmp1[i] = mkc(two * alpha * IM(a1), 0.0);
        s->tmp2[i] = mkc(two * alpha * IM(a2), 0.0);
    }
    orc_fft2d(N, s->tmp1, s->tmp1, -1);
    orc_fft2d(N, s->tmp2, s->tmp2, -1);
    orc_fft2d(N, s->cmpxtmp, s->cmpxtmp, -1);
    orc_fft2d(N, s->tmp3, s->tmp3, -1);
    const int ns = p->nshell;
    const double enorm = sq(1.0 / ((double)N * (double)N));
    memset(spec, 0, sizeof(double) * 6 * (ns + 1));
    FOR_K(s) {
        size_t k = KIDX(s);
        double kx = K1, ky = K2;
        int mshl = orc_shell(K1, K2);
        double kr2 = kx * kx + ky * ky;
        orc_c aa1 = s->tmp1[k], aa2 = s->tmp2[k];
        orc_c bb1, bb2;
        if (K1 == 0 && K2 == 0) { bb1 = 0; bb2 = 0; }
        else {
            bb1 = cmul(mkc(0.0, -(kx / kr2)), aa1);
            bb2 = cmul(mkc(0.0, -(ky / kr2)), aa2);
        }
        orc_c bb = cadd(bb1, bb2);
        orc_c cc1 = cmul(mkc(0.0, kx), bb), cc2 = cmul(mkc(0.0, ky), bb);
        orc_c dd1 = csub(aa1, cc1), dd2 = csub(aa2, cc2);
        orc_c ee1 = s->cmpxtmp[k], ee2 = s->tmp3[k];
        orc_c ff = s->str1[k], gg = s->kpsi[k];
        double ekint = enorm * (sq(cabs_f(aa1)) + sq(cabs_f(aa2)));
        double ekinc = enorm * (sq(cabs_f(cc1)) + sq(cabs_f(cc2)));
        double ekini = enorm * (sq(cabs_f(dd1)) + sq(cabs_f(dd2)));
        double equnt = enorm * (sq(cabs_f(ee1)) + sq(cabs_f(ee2)));
        double eint = enorm * sq(cabs_f(ff));
        double eocc = enorm * sq(cabs_f(gg));
        if (orc_mask(p, K1, K2)) { ekint = ekini = ekinc = equnt = eint = eocc = 0.0; }
        spec[0 * (ns + 1) + mshl] += ekint * half;
        spec[1 * (ns + 1) + mshl] += ekini * half;
        spec[2 * (ns + 1) + mshl] += ekinc * half;
        spec[3 * (ns + 1) + mshl] += equnt * two * (alpha * alpha);
        spec[4 * (ns + 1) + mshl] += eint * half * p->g * two * alpha;
        spec[5 * (ns + 1) + mshl] += eocc;
    }
    scal[0] = egrad; scal[1] = echem; scal[2] = eintr; scal[3] = egrad + eintr;
    scal[4] = norm; scal[5] = momx; scal[6] = momy; scal[7] = spec[5 * (ns + 1) + 0];
    /* per-particle forces (:528-533) */
    if (p->np > 0) {
        double f[2];
        for (int ip = 0; ip < p->np; ++ip) {
            orc_particle_force(s, ip, f);
            s->frc[2 * ip] = f[0];
            s->frc[2 * ip + 1] = f[1];
        }
    }
}
