// Spatio-temporal spectra, resident in HBM.
//
// The reference appends k-space snapshots to files during the run (sts_gpe,
// src/diagnostics.f90:636-695: kpsi(ky, kx=0) and the kpsi planes) and
// post-processes them with a separate program (Run/CompSTS/driver.f90:
// 134-265): Hann window w_n = (1 - cos(2 pi n/(Nt-1)))/2 over the Nt
// snapshots, forward DFT in time, stsnorm |.|^2 with
// stsnorm = (1/Nt)^2 (1/(Nx Ny))^2, and for the planes a sum over shells
// mshl = nint(|k|).  Here the snapshots of the retained modes (the only
// non-zero ones of a dealiased kpsi) stay on the device, mode-major
// X[m][n] (each mode's time series contiguous, so the time transforms read
// whole cache lines); the transforms are the library's register-resident
// Stockham FFT (one per mode), and the shell sums are fixed-order
// reductions over a shell-sorted mode list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "vikshobha.h"
#include "vk_cfg.hpp"

using vk::cplx;

int vk_set_error(int code, const std::string& msg);   // vk_gpe.hip
cplx* vk_internal_kpsi(vk_gpe* h);
int vk_internal_N(const vk_gpe* h);
hipStream_t vk_internal_stream(const vk_gpe* h);

namespace {

#define SCHK(x)                                                                                          \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return vk_set_error(VK_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Modes per time-transform workgroup.  Fewer modes than the column passes'
// TC give more, smaller workgroups per CU whose load and transform phases
// overlap (each mode's series is contiguous, so no access gets narrower):
// measured at 128^2 x 2048 0.077 -> 0.056 ms with 2 modes, at 64^2 x 4096
// 0.046 -> 0.032 ms with 1, at 256^2 x 512 0.068 -> 0.050 ms with 2.  Short
// series keep TC so that a workgroup still fills a wave.  Also the LDS
// stride between the transform regions (staggered so that interleaved
// transforms start on different banks).
template <class C> constexpr int sts_g() { return C::N >= 4096 ? 1 : (C::N >= 256 ? 2 : C::TC); }
template <class C, int G> constexpr int sts_rs() { return C::LDSW + VK_STAGGER / G; }

// X[m][n] = K[kidx[m]] for snapshot n (each mode's time series contiguous)
__global__ void k_sts_gather(const cplx* __restrict__ K, const int* __restrict__ kidx, int nm, int Nt, int n,
                             cplx* __restrict__ X)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < nm) X[(size_t)m * Nt + n] = K[kidx[m]];
}

// One workgroup: TC modes, each transformed along time (Nt = C::N points,
// contiguous per mode).  P[m][w] = stsnorm |sum_n h_n X[m][n] e^{-2 pi i n w/Nt}|^2
// with the window h from a table.  Blocks past the last mode return before
// the transform's barriers (whole workgroups only).
template <class C, int G>
__global__ void __launch_bounds__(G * C::T) k_sts_time(const cplx* __restrict__ X, int nm, double* __restrict__ P,
                                                      const cplx* __restrict__ tw, const double* __restrict__ hann,
                                                      double stsnorm)
{
    constexpr int Nt = C::N, S = C::S, T = C::T;
    extern __shared__ double lds[];
    if ((int)blockIdx.x * G >= nm) return;
    const int u = threadIdx.x;
    const int c = u % G, t = u / G;
    const int m = blockIdx.x * G + c;
    const bool live = m < nm;
    double* lg = lds + c * sts_rs<C, G>();
    const cplx* xr = X + (size_t)(live ? m : 0) * Nt;
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int n = t + T * s;
        const double h = hann[n];
        const cplx x = live ? xr[n] : vk::mk(0.0, 0.0);
        v[s] = vk::mk(h * x.x, h * x.y);
    }
    vk::fft_fwd<Nt, S, C::SPLIT>(v, lg, t, tw);
    if (live) {
        double* pr = P + (size_t)m * Nt;
#pragma unroll
        for (int s = 0; s < S; ++s) pr[t + T * s] = stsnorm * (v[s].x * v[s].x + v[s].y * v[s].y);
    }
}

// spec2d[w][sh] = sum over the shell's modes of P[m][w].  One workgroup:
// one shell x 64 frequencies; wave j sums modes off[sh]+j, +BW, ... in list
// order and the BW partial sums are added in wave order (a fixed order:
// the result does not depend on scheduling).
constexpr int BW = 4;
__global__ void __launch_bounds__(64 * BW) k_sts_bin(const double* __restrict__ P, const int* __restrict__ off,
                                                   int nsh, int Nt, double* __restrict__ spec)
{
    __shared__ double part[BW][64];
    const int lane = threadIdx.x % 64, j = threadIdx.x / 64;
    const int wblk = (Nt + 63) / 64;
    const int sh = blockIdx.x / wblk;
    const int w = (blockIdx.x % wblk) * 64 + lane;
    const bool live = w < Nt;
    const int m0 = off[sh], m1 = off[sh + 1];
    double acc = 0.0;
    if (live) {
#pragma unroll 4
        for (int m = m0 + j; m < m1; m += BW) acc += P[(size_t)m * Nt + w];
    }
    part[j][lane] = acc;
    __syncthreads();
    if (j == 0 && live) {
        double r = part[0][lane];
#pragma unroll
        for (int q = 1; q < BW; ++q) r += part[q][lane];
        spec[(size_t)w * nsh + sh] = r;
    }
}

// spec1d[w][ky] = P[m0[ky]][w] (kx = 0 modes; zero where ky is not retained)
__global__ void k_sts_x0(const double* __restrict__ P, const int* __restrict__ m0, int N, int Nt,
                         double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * Nt) return;
    const int ky = i % N, w = i / N;
    out[i] = m0[ky] >= 0 ? P[(size_t)m0[ky] * Nt + w] : 0.0;
}

unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

bool masked(long kx, long ky, long N) { return 9 * (kx * kx + ky * ky) >= N * N; }
int ksgn(int i, int N) { return i < N / 2 ? i : i - N; }

}  // namespace

struct vk_sts {
    vk_gpe* h = nullptr;
    int N = 0, nt = 0, nm = 0, nsh = 0, count = 0;
    int *kidx = nullptr, *off = nullptr, *m0 = nullptr;
    cplx *X = nullptr, *tw = nullptr;
    double *P = nullptr, *spec2 = nullptr, *spec1 = nullptr, *hann = nullptr;
};

namespace {

template <class F>
int dispatch_nt(int nt, F&& f)
{
    switch (nt) {
    case 32: return f(std::integral_constant<int, 32>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
    case 256: return f(std::integral_constant<int, 256>{});
    case 512: return f(std::integral_constant<int, 512>{});
    case 1024: return f(std::integral_constant<int, 1024>{});
    case 2048: return f(std::integral_constant<int, 2048>{});
    case 4096: return f(std::integral_constant<int, 4096>{});
    default: return vk_set_error(VK_ERR_UNSUPPORTED, "sts: the number of snapshots must be a power of two in [32, 4096]");
    }
}

// the three compute kernels on stream st; ev (optional, 4 events) brackets them
int launch_compute(vk_sts* s, hipStream_t st, hipEvent_t* ev)
{
    const double nn = (double)s->N * s->N;
    const double stsnorm = (1.0 / ((double)s->nt * s->nt)) * (1.0 / (nn * nn));
    if (ev) SCHK(hipEventRecord(ev[0], st));
    int rc = dispatch_nt(s->nt, [&](auto n) {
        using C = vk::Cfg<decltype(n)::value>;
        constexpr int G = sts_g<C>();
        const size_t lds = (size_t)G * sts_rs<C, G>() * sizeof(double);
        auto kern = k_sts_time<C, G>;
        if (lds > 65536)
            SCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, dim3(nblk(s->nm, G)), dim3(G * C::T), lds, st, (const cplx*)s->X, s->nm,
                           s->P, (const cplx*)s->tw, (const double*)s->hann, stsnorm);
        return 0;
    });
    if (rc) return rc;
    if (ev) SCHK(hipEventRecord(ev[1], st));
    hipLaunchKernelGGL(k_sts_bin, dim3((unsigned)(s->nsh * ((s->nt + 63) / 64))), dim3(64 * BW), 0, st,
                       (const double*)s->P, (const int*)s->off, s->nsh, s->nt, s->spec2);
    if (ev) SCHK(hipEventRecord(ev[2], st));
    hipLaunchKernelGGL(k_sts_x0, dim3(nblk((size_t)s->N * s->nt, 256)), dim3(256), 0, st, (const double*)s->P,
                       (const int*)s->m0, s->N, s->nt, s->spec1);
    if (ev) SCHK(hipEventRecord(ev[3], st));
    SCHK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int vk_sts_create(vk_sts** out, vk_gpe* h, int nt)
{
    if (!out || !h) return vk_set_error(VK_ERR_ARG, "sts_create: null argument");
    *out = nullptr;
    int rc = dispatch_nt(nt, [](auto) { return 0; });
    if (rc) return rc;
    const int N = vk_internal_N(h);
    // retained modes (the dealias rule of src/gpe.f90:346), shell-sorted;
    // within a shell in the reference's loop order (kx index outer, ky inner)
    // mshl = nint(|k|) with |k| < N/3 on the retained disk
    const int nsh = (int)std::lround(N / 3.0) + 1;
    std::vector<std::vector<int>> by(nsh);
    std::vector<int> m0(N, -1);
    for (int i1 = 0; i1 < N; ++i1)
        for (int i2 = 0; i2 < N; ++i2) {
            const int k1 = ksgn(i1, N), k2 = ksgn(i2, N);
            if (masked(k1, k2, N)) continue;
            const int sh = (int)std::lround(std::sqrt((double)(k1 * k1 + k2 * k2)));
            if (sh >= nsh) return vk_set_error(VK_ERR_STATE, "sts_create: shell index out of range");
            by[sh].push_back(i1 * N + i2);
        }
    std::vector<int> kidx, off(nsh + 1, 0);
    for (int sh = 0; sh < nsh; ++sh) {
        off[sh] = (int)kidx.size();
        for (int k : by[sh]) {
            if (k / N == 0) m0[k % N] = (int)kidx.size();
            kidx.push_back(k);
        }
    }
    off[nsh] = (int)kidx.size();
    vk_sts* s = new vk_sts();
    s->h = h, s->N = N, s->nt = nt, s->nm = (int)kidx.size(), s->nsh = nsh;
    auto cleanup = [&](int code) { vk_sts_destroy(s); return code; };
    const size_t nm = s->nm;
    if (hipMalloc(&s->kidx, nm * sizeof(int)) != hipSuccess || hipMalloc(&s->off, off.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&s->m0, N * sizeof(int)) != hipSuccess || hipMalloc(&s->X, (size_t)nt * nm * sizeof(cplx)) != hipSuccess ||
        hipMalloc(&s->P, (size_t)nt * nm * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->spec2, (size_t)nt * nsh * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->spec1, (size_t)nt * N * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->tw, (size_t)nt * sizeof(cplx)) != hipSuccess ||
        hipMalloc(&s->hann, (size_t)nt * sizeof(double)) != hipSuccess)
        return cleanup(vk_set_error(VK_ERR_HIP, "sts_create: device allocation failed"));
    std::vector<cplx> tw(nt);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < nt; ++k) {
        const long double a = -2.0L * pi * (long double)k / (long double)nt;
        tw[k] = vk::cplx{(double)cosl(a), (double)sinl(a)};
    }
    // Hann window (1 - cos(2 pi n/(Nt-1)))/2 (Run/CompSTS/driver.f90:163-165)
    std::vector<double> hann(nt);
    const double two_pi = 2.0 * 3.141592653589793;   // in double, as the reference evaluates it
    for (int n = 0; n < nt; ++n) hann[n] = 0.5 * (1.0 - std::cos(two_pi * n / (double)(nt - 1)));
    if (hipMemcpy(s->kidx, kidx.data(), nm * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->m0, m0.data(), N * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->tw, tw.data(), nt * sizeof(cplx), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->hann, hann.data(), nt * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(vk_set_error(VK_ERR_HIP, "sts_create: upload failed"));
    *out = s;
    return 0;
}

void vk_sts_destroy(vk_sts* s)
{
    if (!s) return;
    void* ps[] = {s->kidx, s->off, s->m0, s->X, s->tw, s->P, s->spec2, s->spec1, s->hann};
    for (void* p : ps)
        if (p) hipFree(p);
    delete s;
}

int vk_sts_count(const vk_sts* s) { return s ? s->count : -1; }
int vk_sts_modes(const vk_sts* s) { return s ? s->nm : -1; }
int vk_sts_shells(const vk_sts* s) { return s ? s->nsh : -1; }

// sts_gpe: one snapshot of the handle's current kpsi (retained modes)
int vk_sts_record(vk_sts* s)
{
    if (!s) return vk_set_error(VK_ERR_ARG, "sts_record: null");
    if (s->count >= s->nt) return vk_set_error(VK_ERR_STATE, "sts_record: all snapshots taken");
    hipStream_t st = vk_internal_stream(s->h);
    hipLaunchKernelGGL(k_sts_gather, dim3(nblk(s->nm, 256)), dim3(256), 0, st, (const cplx*)vk_internal_kpsi(s->h),
                       (const int*)s->kidx, s->nm, s->nt, s->count, s->X);
    SCHK(hipGetLastError());
    s->count++;
    return 0;
}

int vk_sts_reset(vk_sts* s)
{
    if (!s) return vk_set_error(VK_ERR_ARG, "sts_reset: null");
    s->count = 0;
    return 0;
}

// spec2d: nt x vk_sts_shells() doubles, shell fastest (spec_sts2d(0:., 1:Nt)
// of CompSTS); spec1d: nt x N doubles, ky fastest (arr1(1:Ny, 1:Nt)); either
// may be NULL.  Needs all nt snapshots.
int vk_sts_compute(vk_sts* s, double* spec2d, double* spec1d)
{
    if (!s) return vk_set_error(VK_ERR_ARG, "sts_compute: null");
    if (s->count != s->nt) return vk_set_error(VK_ERR_STATE, "sts_compute: needs all nt snapshots");
    hipStream_t st = vk_internal_stream(s->h);
    int rc = launch_compute(s, st, nullptr);
    if (rc) return rc;
    if (spec2d)
        SCHK(hipMemcpyAsync(spec2d, s->spec2, (size_t)s->nt * s->nsh * sizeof(double), hipMemcpyDefault, st));
    if (spec1d) SCHK(hipMemcpyAsync(spec1d, s->spec1, (size_t)s->nt * s->N * sizeof(double), hipMemcpyDefault, st));
    SCHK(hipStreamSynchronize(st));
    return 0;
}

// average duration in ms of the three compute kernels (time transforms,
// shell sums, kx = 0 lines) over reps launches each, from HIP events on the
// handle's stream
int vk_sts_kernel_times(vk_sts* s, int reps, double* ms3)
{
    if (!s || reps <= 0 || !ms3) return vk_set_error(VK_ERR_ARG, "sts_kernel_times: bad arguments");
    if (s->count != s->nt) return vk_set_error(VK_ERR_STATE, "sts_kernel_times: needs all nt snapshots");
    hipStream_t st = vk_internal_stream(s->h);
    hipEvent_t ev[4];
    for (auto& e : ev) SCHK(hipEventCreate(&e));
    double acc[3] = {0.0, 0.0, 0.0};
    int rc = 0;
    for (int r = 0; r < reps && !rc; ++r) {
        rc = launch_compute(s, st, ev);
        if (!rc && hipEventSynchronize(ev[3]) != hipSuccess) rc = vk_set_error(VK_ERR_HIP, "sts_kernel_times: event");
        for (int k = 0; k < 3 && !rc; ++k) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess)
                rc = vk_set_error(VK_ERR_HIP, "sts_kernel_times: elapsed");
            acc[k] += ms;
        }
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k) ms3[k] = acc[k] / reps;
    return 0;
}

// the recorded kpsi(ky, kx = 0) columns, count x N complex (the reference's
// kpsi_x0 file content); zero where ky is not retained
int vk_sts_get_x0(vk_sts* s, double* x0)
{
    if (!s || !x0) return vk_set_error(VK_ERR_ARG, "sts_get_x0: null");
    hipStream_t st = vk_internal_stream(s->h);
    SCHK(hipStreamSynchronize(st));
    std::vector<int> m0(s->N);
    SCHK(hipMemcpy(m0.data(), s->m0, s->N * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<cplx> series(s->count > 0 ? s->count : 1);
    std::fill(x0, x0 + 2 * (size_t)s->count * s->N, 0.0);
    for (int ky = 0; ky < s->N; ++ky) {
        if (m0[ky] < 0 || s->count == 0) continue;
        SCHK(hipMemcpy(series.data(), s->X + (size_t)m0[ky] * s->nt, s->count * sizeof(cplx), hipMemcpyDeviceToHost));
        for (int n = 0; n < s->count; ++n) {
            x0[2 * ((size_t)n * s->N + ky)] = series[n].x;
            x0[2 * ((size_t)n * s->N + ky) + 1] = series[n].y;
        }
    }
    return 0;
}

// snapshot n as a full kpsi_t plane (kx*N + ky, zero off the retained
// modes): the content of the reference's kpsi_nrst<n>p<r>.dat records
int vk_sts_get_kpsi(vk_sts* s, int n, double* kpsi_t)
{
    if (!s || !kpsi_t) return vk_set_error(VK_ERR_ARG, "sts_get_kpsi: null");
    if (n < 0 || n >= s->count) return vk_set_error(VK_ERR_ARG, "sts_get_kpsi: snapshot not recorded");
    hipStream_t st = vk_internal_stream(s->h);
    SCHK(hipStreamSynchronize(st));
    std::vector<int> kidx(s->nm);
    std::vector<cplx> col(s->nm);
    SCHK(hipMemcpy(kidx.data(), s->kidx, s->nm * sizeof(int), hipMemcpyDeviceToHost));
    SCHK(hipMemcpy2D(col.data(), sizeof(cplx), s->X + n, (size_t)s->nt * sizeof(cplx), sizeof(cplx), s->nm,
                     hipMemcpyDeviceToHost));
    std::fill(kpsi_t, kpsi_t + 2 * (size_t)s->N * s->N, 0.0);
    for (int m = 0; m < s->nm; ++m) {
        kpsi_t[2 * (size_t)kidx[m]] = col[m].x;
        kpsi_t[2 * (size_t)kidx[m] + 1] = col[m].y;
    }
    return 0;
}

}  // extern "C"
