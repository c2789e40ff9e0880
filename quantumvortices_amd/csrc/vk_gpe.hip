// Host side of the vikshobha library: handle, pipelines, C ABI.
//
// One RK4 stage (src/gpe.f90:357-492 around gpe_rhs :498-671) is four
// passes over HBM:
//   P1 row   : psi(y,.) = IFFT_x Hpsi ; rho = |psi|^2 ; Hrho = FFT_x rho
//   P2 column: rho-hat = FFT_y Hrho ; P_G ; Hrho_d = IFFT_y
//   P3 row   : rho_d, V, psi rows ; N = (g rho_d + V) psi ; HN = FFT_x N
//   P4 column: N-hat = FFT_y HN ; RHS, RK4 accumulation ; Hpsi = IFFT_y kpsi
// kpsi is dealiased, so only |kx| <= N/3 columns are stored and moved;
// rho, rho_d and V are real, so only their kx >= 0 halves are moved.
#include "vk_ens.hpp"
#include "vikshobha.h"

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

using vk::cplx;
using vk::Geo;

// single-field handles of at least this size keep x-space psi rows between
// P1 and P3 (one N^2 field more; P3 then runs two transforms instead of three)
#ifndef VK_XPSI_MIN_N
#define VK_XPSI_MIN_N 2048
#endif

namespace {
thread_local std::string g_err;
}

// Scratch of the allocation-free entry point vk_fft2d_noalloc: one
// half-spectral field (the row pass output the column pass reads) of the
// largest size that entry point accepts, VK_NOALLOC_MAX_N^2 complex doubles
// = 16 MiB.  Every GPU the library runs on holds its own copy until the
// process ends.  Host code never passes this name to a kernel: its device
// address on each device is looked up once, in vk_gpe_create with that
// device current (scratch_setup), and kept in g_scratch_ptr.
#define VK_NOALLOC_MAX_N 1024
#define VK_MAX_DEVICES 64
static __device__ __align__(256) double2 scratch[(size_t)VK_NOALLOC_MAX_N * VK_NOALLOC_MAX_N];

// shared with the other translation units of the library (vk_last_error)
int vk_set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

namespace {

int fail(int code, const std::string& msg) { return vk_set_error(code, msg); }

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(VK_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum PassKind { PK_P1 = 0, PK_P2, PK_P3, PK_P4, PK_PART, PK_ARG4, PK_OTHER, PK_COUNT };
const char* kPassNames[PK_COUNT] = {"p1_rho_row", "p2_dealias_col", "p3_nonlin_row", "p4_rk4_col",
                                    "particles", "argle_update_col", "other"};

int shell_of(int kx, int ky)
{
    // nint(sqrt(kx^2+ky^2)) (src/diagnostics.f90:341-343) by the integer
    // identity  m(m-1) < s <= m(m+1)
    const long long s = (long long)kx * kx + (long long)ky * ky;
    if (s == 0) return 0;
    long long m = (long long)std::sqrt((double)s);
    while (m * (m + 1) < s) ++m;
    while (m > 0 && (m - 1) * m >= s) --m;
    return (int)m;
}

bool masked_host(int kx, int ky, int N) { return 9LL * ((long long)kx * kx + (long long)ky * ky) >= (long long)N * N; }

std::mutex g_scratch_mu;
cplx* g_scratch_ptr[VK_MAX_DEVICES] = {};

// device address of `scratch` on device dev, which must be current; looked
// up on the first handle created on that device and kept
int scratch_setup(int dev)
{
    if (dev < 0 || dev >= VK_MAX_DEVICES) return fail(VK_ERR_UNSUPPORTED, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(VK_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<cplx*>(ptr);
    return 0;
}
int ksg_host(int i, int N) { return i < N / 2 ? i : i - N; }

}  // namespace

struct vk_gpe {
    vk_gpe_params p{};
    int N = 0;
    int nshell = 0;
    Geo geo{};
    int NP = 0, TC = 0;
    hipStream_t stream = nullptr;
    hipStream_t pstream = nullptr;     // particle update, overlapped with P4/P1/P2
    hipEvent_t ev_p3 = nullptr, ev_ps = nullptr;
    cplx* tw = nullptr;
    // state and RK4/ARGLE work fields
    cplx *K0 = nullptr, *K1 = nullptr, *K2 = nullptr, *Hpsi = nullptr, *Hrho = nullptr, *HN = nullptr;
    cplx* Xs = nullptr;   // x-space psi rows written by P1 and read by P3 (xpsi)
    int xpsi = 0;
    int dev = 0;        // device the handle was created on
    int ncu = 0;        // compute units of the device
    int p1pf = 0;       // RK4 P1 as the persistent prefetching kernel (4096^2)
    int tr1_p1 = 0, tr1_p3 = 0;   // RK4 P1 / P3 (psi-row path) with one row per workgroup
    // advective ARGLE (Taylor-Green and/or counterflow); Hdy = IFFT_y(i ky kpsi)
    int adv_tg = 0;
    double adv_vx = 0.0, adv_vy = 0.0;
    cplx* Hdy = nullptr;
    // caller-supplied advection velocity (replaces the Taylor-Green field)
    double *Ux = nullptr, *Uy = nullptr;
    int adv_field = 0;
    // SGLE (GPE_SGLE): noise field (x-space, then k-space), the chemical
    // potential on the device, norm partials, the draw counter of the
    // generator
    int sgle = 0;
    double sg_beta = 0.0, sg_nuN = 0.0, sg_ktrunc = 0.0;
    unsigned long long sg_seed = 0, sg_draw = 0;
    cplx *Xn = nullptr, *Kn = nullptr;
    double *omega_dev = nullptr, *npart = nullptr;
    cplx *HV = nullptr, *KV = nullptr, *Krho = nullptr, *Vhat = nullptr, *EX = nullptr, *EY = nullptr;
    double *fpart = nullptr, *pst = nullptr, *part = nullptr, *dsum = nullptr;
    unsigned char* phit = nullptr;   // real-space path: window flags per row block and particle
    int hstate = 0;   // 0: Hpsi stale, 1: Hpsi = IFFT_y K0 on all panels, 2: K0 dealiased, Hpsi on retained panels,
                      // 3: K0 dealiased, Hpsi on the retained panels of the single-field shape svar (layout CfgE<N, svar>)
    int kv_valid = 0;
    int have_part = 0;
    int np_cap = 0;     // particle buffers are sized for this many particles
    int rs_part = 0;   // particle potential and force evaluated in real space inside P3
    // diagnostics
    cplx* X[7] = {};
    cplx* Kd[5] = {};
    cplx* Hd = nullptr;
    int *modes = nullptr, *offs = nullptr;
    double *dspec = nullptr, *dpart = nullptr;
    int diag_ready = 0;
    // ensemble shape variant of this member (CfgE<N, ens_var>) and its
    // panel geometry; Hpsi holds that variant's layout only between the
    // ensemble's preparation and the member's next single-field access
    int ens_var = 0;
    Geo geo_v{};
    // one-member ensemble table: N <= 1024 runs the same kernels as a
    // batched ensemble member, so both give bit-identical results; svar is
    // its transform shape (0 as a default ensemble member, 1: CfgE<N, 1>)
    int svar = 0;
    Geo geo_s{};
    vk::EnsReal* e1R = nullptr;
    int* e1map[3] = {};
    int* e1off[3] = {};
    int e1nblk[3] = {};
    size_t e1lds[4] = {};
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> evkind;
    size_t evn = 0;
};

namespace {

size_t f2(const vk_gpe* h) { return (size_t)h->N * h->N; }

template <int N> using CfgN = vk::Cfg<N>;
// The same panel layout with one row per row-pass workgroup (256 threads)
// instead of TR rows: more, smaller workgroups per CU at the cost of half
// as long row pieces of the panels.
template <class C> using CfgTR1 = vk::CfgT<C::N, C::S, C::TC, 1, C::SPLIT>;

template <class F>
int dispatch(int N, F&& f)
{
    switch (N) {
    case 32: return f(std::integral_constant<int, 32>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
    case 256: return f(std::integral_constant<int, 256>{});
    case 512: return f(std::integral_constant<int, 512>{});
    case 1024: return f(std::integral_constant<int, 1024>{});
    case 2048: return f(std::integral_constant<int, 2048>{});
    case 4096: return f(std::integral_constant<int, 4096>{});
    default: return fail(VK_ERR_UNSUPPORTED, "N must be a power of two in [32, 4096]");
    }
}

void ev_begin(vk_gpe* h)
{
    if (!h->timing) return;
    if (h->evn + 2 > h->ev.size()) {
        for (int i = 0; i < 64; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) == hipSuccess) h->ev.push_back(e);
        }
    }
    hipEventRecord(h->ev[h->evn], h->stream);
}
void ev_end(vk_gpe* h, int kind)
{
    if (!h->timing) return;
    hipEventRecord(h->ev[h->evn + 1], h->stream);
    h->evkind.push_back(kind);
    h->evn += 2;
}

std::mutex g_lds_mu;
std::map<const void*, size_t> g_lds_max;   // dynamic-LDS limit set per kernel so far

template <class Kern>
void set_lds(Kern k, size_t bytes)
{
    // the limit only grows: a kernel whose LDS depends on a runtime size
    // (P3 with np particles) may be launched larger than at its first use
    if (bytes <= 65536) return;
    std::lock_guard<std::mutex> lk(g_lds_mu);
    size_t& cur = g_lds_max[(const void*)k];
    if (bytes > cur) {
        hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        cur = bytes;
    }
}

template <class C, class Kern, class... A>
void row_launch_x(vk_gpe* h, int kind, size_t extra, Kern k, A... args)
{
    const size_t lds = (size_t)C::TR * C::ROWRS * sizeof(double) + extra;
    set_lds(k, lds);
    ev_begin(h);
    hipLaunchKernelGGL(k, dim3(C::N / C::TR), dim3(C::ROWT), lds, h->stream, args...);
    ev_end(h, kind);
}

template <class C, class Kern, class... A>
void row_launch(vk_gpe* h, int kind, Kern k, A... args)
{
    row_launch_x<C>(h, kind, 0, k, args...);
}

// P3 keeps S doubles per thread in LDS besides the transform regions (in
// registers instead with xpsi and no real-space particles), then the
// particle positions and a one-bit-per-particle hit mask
template <class C> size_t p3_extra(const vk_gpe* h)
{
    const bool park = !h->xpsi || (h->have_part && h->p.np > 0 && h->rs_part && !vk::kP3FregPart);
    const size_t np = (size_t)std::max(h->p.np, 1);
    return (park ? (size_t)C::S * C::ROWT * sizeof(double) : 0) + 2 * np * sizeof(double);
}

template <class C, class Kern, class... A>
void col_launch(vk_gpe* h, int kind, int nblocks, Kern k, A... args)
{
    const size_t lds = (size_t)C::TC * C::COLRS * sizeof(double);
    set_lds(k, lds);
    ev_begin(h);
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(C::COLT), lds, h->stream, args...);
    ev_end(h, kind);
}

// P2 columns per workgroup.  P2 covers only the kx >= 0 panels.  At
// 4096^2 one column per workgroup (1366 workgroups) is the fastest: with
// 2-column panels 4% faster than whole panels, with 4-column panels
// 0.0697 against 0.0707 ms for two columns; at 2048^2 one column is 10%
// slower than whole panels.  Results are bitwise the same either way.
#ifndef VK_P2_COLS_4096
#define VK_P2_COLS_4096 1
#endif
template <class C> constexpr int p2_cols() { return C::N == 4096 ? VK_P2_COLS_4096 : C::TC; }
// P4 columns per workgroup at 4096^2 (0: whole panels).  Two of the four
// columns of a panel in a 512-thread workgroup: 0.170 ms against 0.189 for
// whole panels (1024 threads, one workgroup per CU) and 0.193 for one.
#ifndef VK_P4_COLS_4096
#define VK_P4_COLS_4096 2
#endif
template <class C> constexpr int p4_cols() { return C::N == 4096 && VK_P4_COLS_4096 > 0 ? VK_P4_COLS_4096 : C::TC; }

// column pass with G columns per workgroup over npan panels
template <class C, int G, class Kern, class... A>
void colg_launch(vk_gpe* h, int kind, int npan, Kern k, A... args)
{
    constexpr int GS = C::TC / G;
    const size_t lds = (size_t)G * C::COLRS * sizeof(double);
    set_lds(k, lds);
    const int grid = GS == 1 ? npan : (npan + 7) / 8 * 8 * GS;
    ev_begin(h);
    hipLaunchKernelGGL(k, dim3(grid), dim3(G * C::T), lds, h->stream, args...);
    ev_end(h, kind);
}

// the flags (last argument: 0, 1 or 3) select a compile-time instantiation
template <class C>
void p2_launch(vk_gpe* h, cplx* Hrho, cplx* HV, const cplx* KV, cplx* Krho, const cplx* tw, Geo g, int flags)
{
    constexpr int G = p2_cols<C>();
    auto go = [&](auto k) { colg_launch<C, G>(h, PK_P2, h->geo.Pp, k, Hrho, HV, KV, Krho, tw, g, flags); };
    if (flags == 0) go(vk::k_p2<C, G, 0>);
    else if (flags == 1) go(vk::k_p2<C, G, 1>);
    else if (flags == 3) go(vk::k_p2<C, G, 3>);
    else go(vk::k_p2<C, G>);
}

// The persistent prefetching P1 (vk::k_p1xp) runs at 4096^2 when the
// runtime panel geometry is the one its LDS slot map was built for and the
// row blocks split evenly over one workgroup per CU in XCD groups of 8.
#ifndef VK_P1_PREFETCH
#define VK_P1_PREFETCH 1
#endif
// the persistent prefetching P1 fits this handle and device (4096^2, a CU
// count whose row blocks divide evenly over the XCDs)
int p1_prefetch_fits(const vk_gpe* h)
{
    if (!VK_P1_PREFETCH || h->N != 4096 || !h->xpsi) return 0;
    using C = CfgN<4096>;
    static_assert(vk::p1p_lds<C>() <= 160 * 1024, "prefetch buffer exceeds LDS");
    const int g8 = h->ncu / 8, r8 = (C::N / C::TR) / 8;
    return h->ncu > 0 && h->ncu % 8 == 0 && r8 % g8 == 0 && h->geo.Pp == vk::geo_pp<C>() &&
           h->geo.Pn == vk::geo_pn<C>();
}

int p1_prefetch_ok(const vk_gpe* h)
{
    // the plain P1, whose loads are all in flight at once, is faster; the
    // prefetching kernel runs when VK_P1_PREFETCH=1 (tests compare the two)
    const char* on = std::getenv("VK_P1_PREFETCH");
    return on && on[0] == '1' && p1_prefetch_fits(h);
}

template <class C>
void p1p_launch(vk_gpe* h, double sc)
{
    if constexpr (C::N == 4096) {
        auto k = vk::k_p1xp<C>;
        const size_t lds = vk::p1p_lds<C>();
        set_lds(k, lds);
        ev_begin(h);
        hipLaunchKernelGGL(k, dim3(h->ncu), dim3(C::ROWT), lds, h->stream, (const cplx*)h->Hpsi, h->Hrho,
                           (const cplx*)h->tw, h->geo, sc, h->Xs);
        ev_end(h, PK_P1);
    }
}

int ncols(const vk_gpe* h, int pmode)
{
    if (pmode == 0) return h->NP;
    if (pmode == 2) return h->geo.Pp;
    return h->geo.Pp + h->geo.Pn;
}

double scale_of(const vk_gpe* h) { return 1.0 / ((double)h->N * (double)h->N); }

// ---- generic 2D transforms ------------------------------------------------

int fwd2d(vk_gpe* h, const cplx* X, cplx* Hs, cplx* K, int op_row, int op_col)
{
    return dispatch(h->N, [&](auto n) {
        using C = CfgN<decltype(n)::value>;
        row_launch<C>(h, PK_OTHER, vk::k_row_fwd<C>, X, Hs, (const cplx*)h->tw, h->geo, 0, op_row);
        constexpr int G = p4_cols<C>();
        colg_launch<C, G>(h, PK_OTHER, h->NP, vk::k_col_fwd<C, G>, (const cplx*)Hs, K, (const cplx*)h->tw, h->geo, 0,
                          op_col);
        return 0;
    });
}

// forward 2D transform of a real field: only the kx >= 0 panels are stored
// by the row pass and transformed by the column pass (the rest follows from
// Hermitian symmetry)
int fwd2d_half(vk_gpe* h, const cplx* X, cplx* Hs, cplx* K, int op_row, int op_col)
{
    return dispatch(h->N, [&](auto n) {
        using C = CfgN<decltype(n)::value>;
        row_launch<C>(h, PK_OTHER, vk::k_row_fwd<C>, X, Hs, (const cplx*)h->tw, h->geo, 2, op_row);
        constexpr int G = p4_cols<C>();
        colg_launch<C, G>(h, PK_OTHER, h->geo.Pp, vk::k_col_fwd<C, G>, (const cplx*)Hs, K, (const cplx*)h->tw, h->geo, 2,
                          op_col);
        return 0;
    });
}

int col_inv(vk_gpe* h, const cplx* K, cplx* Hs, int pmode, int op)
{
    return dispatch(h->N, [&](auto n) {
        using C = CfgN<decltype(n)::value>;
        constexpr int G = p4_cols<C>();
        colg_launch<C, G>(h, PK_OTHER, ncols(h, pmode), vk::k_col_inv<C, G>, K, Hs, (const cplx*)h->tw, h->geo, pmode,
                          op);
        return 0;
    });
}

int row_inv(vk_gpe* h, const cplx* Hs, cplx* X, int pload, int op, double scale)
{
    return dispatch(h->N, [&](auto n) {
        using C = CfgN<decltype(n)::value>;
        row_launch<C>(h, PK_OTHER, vk::k_row_inv<C>, Hs, X, (const cplx*)h->tw, h->geo, pload, op, scale);
        return 0;
    });
}

int dealias_k(vk_gpe* h, cplx* K)
{
    hipLaunchKernelGGL(vk::k_dealias, dim3(1024), dim3(256), 0, h->stream, K, h->N);
    return 0;
}

// ---- state management -------------------------------------------------------

int ensure_any(vk_gpe* h)
{
    if (h->hstate == 0 || h->hstate == 3) {
        int rc = col_inv(h, h->K0, h->Hpsi, 0, 0);
        if (rc) return rc;
        h->hstate = 1;
    }
    return 0;
}

int ensure_rk4(vk_gpe* h)
{
    if (h->hstate != 2) {
        dealias_k(h, h->K0);
        int rc = col_inv(h, h->K0, h->Hpsi, 1, 0);
        if (rc) return rc;
        h->hstate = 2;
    }
    return 0;
}

int ensure_full(vk_gpe* h)
{
    if (h->hstate != 1) {
        int rc = col_inv(h, h->K0, h->Hpsi, 0, 0);
        if (rc) return rc;
        h->hstate = 1;
    }
    return 0;
}

// ---- particles ----------------------------------------------------------------

double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

vk::PartArgs part_args(const vk_gpe* h, int nrows, int half, int stage, int src = 0, int nblk = 0)
{
    const double pi = 4.0 * std::atan(1.0);
    const double L = 2.0 * pi;
    const double dx = L / h->N;
    vk::PartArgs a{};
    a.np = h->p.np;
    a.N = h->N;
    a.nrows = nrows;
    a.half = half;
    a.coef = 2.0 * h->p.alpha * (dx * dx) / ((double)h->N * (double)h->N);
    const double dE = (double)0.062f;         // src/gpe_particle.f90:395, default-real literal
    const double r0 = 4.0 * h->p.xi;         // :396
    a.rf = 384.0 * dE * ipow(pi, 12) * ipow(r0, 12) / ipow(L, 12);
    a.efac = 64.0 * dE * ipow(pi, 12) * ipow(r0, 12) / ipow(L, 12);
    a.dt = h->p.dt;
    a.mass = h->p.mass;
    a.stage = stage;
    a.src = src;
    a.nblk = nblk;
    a.coef_rs = 2.0 * h->p.alpha * (dx * dx) / ((double)h->N * (double)h->N);
    return a;
}

// Window radius (in the minimum-image displacement a) of the real-space
// particle potential.  The potential depends on the periodic coordinate
// X(a) = 2(s + s^3/6 + s^5/40), s = sin(a/2), which saturates at
// X(pi) = 2.383 near the half box, so V at the edge a = 10 d is
// V0 exp(-X(10 d)^2/(2 d^2)), not V0 exp(-50).  The 10 d window is kept when
// that is below exp(-45) (3e-20) of the peak (X(a) = a - O(a^5) for small
// a, so narrow potentials keep it); otherwise the window covers the whole
// period (any radius >= pi does).
double real_space_window(double d)
{
    const double pi = 4.0 * std::atan(1.0);
    const double a = std::min(10.0 * d, pi);
    const double s = std::sin(0.5 * a), s2 = s * s;
    const double X = 2.0 * (s + (1.0 / 6.0) * (s2 * s) + (1.0 / 40.0) * (s2 * s2 * s));
    if (10.0 * d < pi && X * X / (2.0 * d * d) >= 45.0) return 10.0 * d;
    return 4.0;
}

vk::P3Part p3part(const vk_gpe* h)
{
    vk::P3Part q{};
    if (h->p.np > 0 && h->have_part) {
        const double pi = 4.0 * std::atan(1.0);
        q.pos = h->pst;
        q.ppart = h->fpart;
        q.phit = h->phit;
        q.np = h->p.np;
        q.V0 = h->p.V0;
        q.d = h->p.d;
        q.W = real_space_window(h->p.d);
        q.dx = 2.0 * pi / h->N;
    }
    return q;
}

// Real-space particle path in P3: valid when the Gaussian's spectrum beyond
// the dealias radius, exp(-(N/3 d)^2 / 2), is below 1e-20 of its peak (and
// the aliasing beyond N/2 smaller still), so the reference's spectral
// translation equals the analytic V(x - q) to far below FP64 roundoff.
int use_real_space(const vk_gpe_params& p)
{
    if (p.np <= 0 || p.d <= 0.0) return 0;
    const double kd = (p.N / 3.0) * p.d;
    return 0.5 * kd * kd > 46.0;
}

// KV on rows [0, nrows) from the current positions
int particle_kv(vk_gpe* h, int nrows)
{
    const int np = h->p.np;
    ev_begin(h);
    hipLaunchKernelGGL(vk::k_ptab, dim3((np * h->N + 255) / 256), dim3(256), 0, h->stream, (const double*)h->pst,
                       h->EX, h->EY, np, h->N);
    hipLaunchKernelGGL(vk::k_kv, dim3(nrows), dim3(256), (size_t)np * sizeof(cplx), h->stream,
                       (const cplx*)h->Vhat, (const cplx*)h->EX, (const cplx*)h->EY, h->KV, np, h->N, nrows);
    ev_end(h, PK_PART);
    return 0;
}

int particle_forces(vk_gpe* h, const cplx* Krho, int nrows, int half, int stage)
{
    const int np = h->p.np;
    constexpr int PG = 8;
    ev_begin(h);
    hipLaunchKernelGGL(vk::k_force_rows<PG>, dim3(nrows, (np + PG - 1) / PG), dim3(256), 0, h->stream, Krho,
                       (const cplx*)h->Vhat, (const cplx*)h->EY, h->fpart, np, h->N);
    hipLaunchKernelGGL(vk::k_pstep, dim3(1), dim3(vk::PSTEP_T), 0, h->stream, (const double*)h->fpart, (const cplx*)h->EX,
                       h->pst, part_args(h, nrows, half, stage));
    ev_end(h, PK_PART);
    return 0;
}

// particle stage update from the real-space block partials written by P3
// The update depends on P3 (force partials) and is needed by the next
// P3 (positions); it runs on a side stream while P4, P1, P2 proceed.
int particle_step_rs(vk_gpe* h, int nblk, int nw, int stage)
{
    const bool side = !h->timing && h->pstream;
    hipStream_t st = side ? h->pstream : h->stream;
    if (side) {
        hipEventRecord(h->ev_p3, h->stream);
        hipStreamWaitEvent(h->pstream, h->ev_p3, 0);
    }
    ev_begin(h);
    const double dx = 2.0 * (4.0 * std::atan(1.0)) / h->N;
    const double coef = 2.0 * h->p.alpha * (dx * dx) / ((double)h->N * (double)h->N);
    hipLaunchKernelGGL(vk::k_pforce_rs, dim3(h->p.np), dim3(256), 0, st, (const double*)h->fpart,
                       (const unsigned char*)h->phit, nblk, nw, h->p.np, coef, h->pst + 28 * h->p.np + 8);
    hipLaunchKernelGGL(vk::k_pstep, dim3(1), dim3(vk::PSTEP_T), 0, st, (const double*)h->fpart, (const cplx*)h->EX,
                       h->pst, part_args(h, 0, 0, stage, 1, nblk));
    ev_end(h, PK_PART);
    if (side) hipEventRecord(h->ev_ps, h->pstream);
    return 0;
}

// order the main stream after the last side-stream particle update
void join_particles(vk_gpe* h)
{
    if (h->pstream && h->ev_ps) hipStreamWaitEvent(h->stream, h->ev_ps, 0);
}

// ---- RK4 / ARGLE -----------------------------------------------------------------

int ens1_step(vk_gpe* h);
int make_single_table(vk_gpe* h);
vk::EnsReal ens_entry(const vk_gpe* h);
int prep_rk4(vk_gpe* h);

int rk4_step(vk_gpe* h)
{
    if (h->e1R && !(h->have_part && h->p.np > 0)) return ens1_step(h);
    return dispatch(h->N, [&](auto n) {
        using C = CfgN<decltype(n)::value>;
        const double sc = scale_of(h);
        const double dt = h->p.dt, oney6 = 1.0 / 6.0;
        const int parts = h->have_part && h->p.np > 0;
        const int rs = parts && h->rs_part;
        const int Kp = C::TC * h->geo.Pp;
        const vk::P3Part pa = p3part(h);
        vk::RKArgs a{};
        a.pref = make_double2(-h->p.disp, -1.0);   // (1 - i disp)/i
        a.alpha = h->p.alpha;
        a.omega = h->p.omega;
        for (int stage = 1; stage <= 4; ++stage) {
            if (parts && !rs) particle_kv(h, Kp);
            if (h->p1pf)
                p1p_launch<C>(h, sc);
            else if (h->xpsi && h->tr1_p1)
                row_launch<CfgTR1<C>>(h, PK_P1, vk::k_p1x<CfgTR1<C>>, (const cplx*)h->Hpsi, h->Hrho, (const cplx*)h->tw,
                                      h->geo, 1, sc, h->Xs);
            else if (h->xpsi)
                row_launch<C>(h, PK_P1, vk::k_p1x<C>, (const cplx*)h->Hpsi, h->Hrho, (const cplx*)h->tw, h->geo, 1, sc,
                              h->Xs);
            else
                row_launch<C>(h, PK_P1, vk::k_p1<C>, (const cplx*)h->Hpsi, h->Hrho, (const cplx*)h->tw, h->geo, 1, sc);
            p2_launch<C>(h, h->Hrho, h->HV, (const cplx*)h->KV, h->Krho,
                          (const cplx*)h->tw, h->geo, (parts && !rs) ? 3 : 0);
            if (parts && !rs) particle_forces(h, h->Krho, Kp, 1, stage);
            if (rs) join_particles(h);   // positions from the previous stage
            if (h->xpsi && h->tr1_p3) {
                using C1 = CfgTR1<C>;
                auto k3 = !parts ? vk::k_p3x<C1, 0> : (rs ? vk::k_p3x<C1, 6> : vk::k_p3x<C1, 1>);
                row_launch_x<C1>(h, PK_P3, p3_extra<C1>(h), k3, (const cplx*)h->Hrho, (const cplx*)h->HV,
                                 (const cplx*)h->Xs, h->HN, (const cplx*)h->tw, h->geo, 1, h->p.g * sc, sc, 1.0, pa);
            } else if (h->xpsi) {
                auto k3 = !parts ? vk::k_p3x<C, 0> : (rs ? vk::k_p3x<C, 6> : vk::k_p3x<C, 1>);
                row_launch_x<C>(h, PK_P3, p3_extra<C>(h), k3, (const cplx*)h->Hrho, (const cplx*)h->HV,
                                (const cplx*)h->Xs, h->HN, (const cplx*)h->tw, h->geo, 1, h->p.g * sc, sc, 1.0, pa);
            } else {
                auto k3 = !parts ? vk::k_p3<C, 0> : (rs ? vk::k_p3<C, 6> : vk::k_p3<C, 1>);
                row_launch_x<C>(h, PK_P3, p3_extra<C>(h), k3, (const cplx*)h->Hrho, (const cplx*)h->HV, (const cplx*)h->Hpsi,
                              h->HN, (const cplx*)h->tw, h->geo, 1, 1, h->p.g * sc, sc, sc, 1.0, pa);
            }
            if (rs) {
                if (h->xpsi && h->tr1_p3) particle_step_rs(h, C::N, (CfgTR1<C>::ROWT + 63) / 64, stage);
                else particle_step_rs(h, C::N / C::TR, (C::ROWT + 63) / 64, stage);
            }
            a.stage = stage;
            a.wacc = (stage == 1 || stage == 4) ? dt * oney6 : dt * oney6 * 2.0;
            a.cstep = stage == 3 ? dt : dt * 0.5;
            a.r1 = (dt * oney6) / (dt * 0.5);
            a.r3 = (dt * oney6 * 2.0) / dt;
            constexpr int G4 = p4_cols<C>();
            auto k4 = stage == 1 ? vk::k_p4<C, 1, G4> : stage == 2 ? vk::k_p4<C, 2, G4>
                                                      : stage == 3 ? vk::k_p4<C, 3, G4> : vk::k_p4<C, 4, G4>;
            colg_launch<C, G4>(h, PK_P4, h->geo.Pp + h->geo.Pn, k4, (const cplx*)h->HN, h->Hpsi, h->K0, h->K1,
                               h->K2, (const cplx*)h->tw, h->geo, a);
        }
        if (rs) join_particles(h);
        h->kv_valid = 0;
        return 0;
    });
}

// SGLE part of one ARGLE step, before the passes: the noise field of this
// step (noise = caller's standard normals, or NULL for the generator's next
// draw) scaled by noisesigma/sqrt(dvol) (src/gpe.f90:150-164) and
// transformed to k-space, and the chemical potential of this step from the
// norm of kpsi at the step's start (finiteT_adhoc_chempot, :226-266).
constexpr int kNormBlocks = 1024;
int sgle_prepare(vk_gpe* h, const double* noise)
{
    const double pi = 4.0 * std::atan(1.0);
    const double L = 2.0 * pi, dx = L / h->N, dvol = dx * dx;
    const double sigma = std::sqrt((1.0 / (2.0 * h->p.alpha)) * h->sg_beta);
    if (noise) HIPCHK(hipMemcpyAsync(h->Xn, noise, f2(h) * sizeof(cplx), hipMemcpyDefault, h->stream));
    hipLaunchKernelGGL(vk::k_sgle_noise, dim3(1024), dim3(256), 0, h->stream, h->Xn, h->N, h->sg_draw, h->sg_seed,
                       1.0 / std::sqrt(dvol), sigma, noise ? 0 : 1);
    if (!noise) h->sg_draw++;
    int rc = fwd2d(h, h->Xn, h->Hd, h->Kn, 0, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(vk::k_knorm, dim3(kNormBlocks), dim3(256), 0, h->stream, (const cplx*)h->K0, f2(h), h->npart);
    hipLaunchKernelGGL(vk::k_chempot, dim3(1), dim3(256), 0, h->stream, (const double*)h->npart, kNormBlocks,
                       h->omega_dev, dvol / ((double)h->N * (double)h->N), h->p.dt, h->sg_nuN, L * L);
    return 0;
}

int argle_step(vk_gpe* h, int want_res, const double* noise = nullptr)
{
    if (h->sgle) {
        int rc = sgle_prepare(h, noise);
        if (rc) return rc;
    }
    return dispatch(h->N, [&](auto n) {
        using C = CfgN<decltype(n)::value>;
        const double sc = scale_of(h);
        const int parts = h->have_part && h->p.np > 0;
        const int rs = parts && h->rs_part;
        const vk::P3Part pa = p3part(h);
        if (parts && !rs && !h->kv_valid) {
            particle_kv(h, C::TC * h->geo.Pp);
            h->kv_valid = 1;
        }
        const bool adv = h->adv_tg || h->adv_field || h->adv_vx != 0.0 || h->adv_vy != 0.0;
        constexpr int GC = p4_cols<C>();
        if (adv)   // i ky kpsi -> Hdy (all columns)
            colg_launch<C, GC>(h, PK_OTHER, h->NP, vk::k_col_inv<C, GC>, (const cplx*)h->K0, h->Hdy,
                               (const cplx*)h->tw, h->geo, 0, 2);
        if (h->xpsi)
            row_launch<C>(h, PK_P1, vk::k_p1x<C>, (const cplx*)h->Hpsi, h->Hrho, (const cplx*)h->tw, h->geo, 0, sc,
                          h->Xs);
        else
            row_launch<C>(h, PK_P1, vk::k_p1<C>, (const cplx*)h->Hpsi, h->Hrho, (const cplx*)h->tw, h->geo, 0, sc);
        p2_launch<C>(h, h->Hrho, h->HV, (const cplx*)h->KV, h->Krho,
                      (const cplx*)h->tw, h->geo, (parts && !rs) ? 1 : 0);
        if (adv) {
            const vk::AdvArgs aa{h->adv_tg, h->adv_vx, h->adv_vy, 0.25 / h->p.alpha, sc,
                                 h->adv_field ? (const double*)h->Ux : nullptr,
                                 h->adv_field ? (const double*)h->Uy : nullptr};
            const cplx* xs = h->xpsi ? (const cplx*)h->Xs : nullptr;
            auto launch = [&](auto k3) {
                row_launch_x<C>(h, PK_P3, p3_extra<C>(h), k3, (const cplx*)h->Hrho, (const cplx*)h->HV,
                                (const cplx*)h->Hpsi, xs, (const cplx*)h->Hdy, h->HN, (const cplx*)h->tw, h->geo,
                                -h->p.g * sc, -sc, sc, -1.0, pa, aa);
            };
            if (h->xpsi) launch(!parts ? vk::k_p3a<C, 0, true> : (rs ? vk::k_p3a<C, 2, true> : vk::k_p3a<C, 1, true>));
            else launch(!parts ? vk::k_p3a<C, 0, false> : (rs ? vk::k_p3a<C, 2, false> : vk::k_p3a<C, 1, false>));
        } else if (h->xpsi) {
            auto k3 = !parts ? vk::k_p3x<C, 0> : (rs ? vk::k_p3x<C, 2> : vk::k_p3x<C, 1>);
            row_launch_x<C>(h, PK_P3, p3_extra<C>(h), k3, (const cplx*)h->Hrho, (const cplx*)h->HV, (const cplx*)h->Xs,
                            h->HN, (const cplx*)h->tw, h->geo, 0, -h->p.g * sc, -sc, -1.0, pa);
        } else {
            auto k3 = !parts ? vk::k_p3<C, 0> : (rs ? vk::k_p3<C, 2> : vk::k_p3<C, 1>);
            row_launch_x<C>(h, PK_P3, p3_extra<C>(h), k3, (const cplx*)h->Hrho, (const cplx*)h->HV, (const cplx*)h->Hpsi, h->HN,
                          (const cplx*)h->tw, h->geo, 0, 0, -h->p.g * sc, -sc, sc, -1.0, pa);
        }
        vk::ArgleArgs a{h->p.alpha, h->p.omega, h->p.dt, want_res};
        if (h->sgle) {
            a.Kn = h->Kn;
            a.omega_dev = h->omega_dev;
            a.sdt = std::sqrt(h->p.dt);
            a.ktrunc2 = h->sg_ktrunc > 0.0 ? h->sg_ktrunc * h->sg_ktrunc : 0.0;
        }
        // the columns per workgroup of P4 (two of four at 4096^2)
        constexpr int GA = p4_cols<C>();
        static_assert(GA == C::TC || C::NP % 8 == 0, "column groups need whole groups of 8 panels");
        colg_launch<C, GA>(h, PK_ARG4, h->NP, vk::k_a4<C, GA>, (const cplx*)h->HN, h->Hpsi, h->K0,
                           (const cplx*)h->tw, a, h->part);
        if (want_res)
            hipLaunchKernelGGL(vk::k_reduce_fixed, dim3(1), dim3(256), 0, h->stream, (const double*)h->part,
                               h->NP * (C::TC / GA), 2, h->dsum);
        return 0;
    });
}

// ---- diagnostics ---------------------------------------------------------------

int diag_setup(vk_gpe* h)
{
    if (h->diag_ready) return 0;
    const size_t fb = f2(h) * sizeof(cplx);
    for (auto& x : h->X) HIPCHK(hipMalloc(&x, fb));
    for (auto& k : h->Kd) HIPCHK(hipMalloc(&k, fb));
    HIPCHK(hipMalloc(&h->Hd, fb));
    HIPCHK(hipMemset(h->Hd, 0, fb));
    // retained modes grouped by shell, in the reference's loop order
    // (kx index outer, ky index inner; src/diagnostics.f90:332-402)
    const int N = h->N, ns = h->nshell;
    std::vector<int> cnt(ns + 2, 0);
    for (int i1 = 0; i1 < N; ++i1)
        for (int i2 = 0; i2 < N; ++i2) {
            const int kx = ksg_host(i1, N), ky = ksg_host(i2, N);
            if (masked_host(kx, ky, N)) continue;
            cnt[shell_of(kx, ky) + 1]++;
        }
    for (int m = 0; m <= ns; ++m) cnt[m + 1] += cnt[m];
    std::vector<int> modes(cnt[ns + 1]);
    std::vector<int> pos(cnt.begin(), cnt.end());
    for (int i1 = 0; i1 < N; ++i1)
        for (int i2 = 0; i2 < N; ++i2) {
            const int kx = ksg_host(i1, N), ky = ksg_host(i2, N);
            if (masked_host(kx, ky, N)) continue;
            modes[pos[shell_of(kx, ky)]++] = i1 * N + i2;
        }
    HIPCHK(hipMalloc(&h->modes, sizeof(int) * std::max<size_t>(1, modes.size())));
    HIPCHK(hipMalloc(&h->offs, sizeof(int) * (ns + 2)));
    HIPCHK(hipMemcpy(h->modes, modes.data(), sizeof(int) * modes.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->offs, cnt.data(), sizeof(int) * (ns + 2), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&h->dspec, sizeof(double) * 6 * (ns + 1)));
    HIPCHK(hipMalloc(&h->dpart, sizeof(double) * 8 * 1024));
    h->diag_ready = 1;
    return 0;
}

// the diagnostics' device work on h->stream: energy sums in h->dsum, the six
// spectra in h->dspec, particle diagnostics in the particle state
int diag_launch(vk_gpe* h)
{
    int rc = diag_setup(h);
    if (rc) return rc;
    if ((rc = ensure_any(h))) return rc;
    const double sc = scale_of(h);
    const int pl = h->hstate == 2 ? 1 : 0;
    cplx **X = h->X, **Kd = h->Kd;
    // psi, d_x psi, d_y psi
    if ((rc = row_inv(h, h->Hpsi, X[0], pl, 0, sc))) return rc;
    if ((rc = row_inv(h, h->Hpsi, X[1], pl, 1, sc))) return rc;
    if ((rc = col_inv(h, h->K0, h->Hd, pl, 2))) return rc;
    if ((rc = row_inv(h, h->Hd, X[2], pl, 0, sc))) return rc;
    // rho-hat (dealiased) -> Kd[4]; P_G rho -> X[3].  rho, P_G rho, V and
    // the w and q fields are real: only their kx >= 0 columns are
    // transformed, the rest is their Hermitian mirror.
    const int Kp = h->TC * h->geo.Pp;
    if ((rc = fwd2d_half(h, X[0], h->Hd, Kd[4], 1, 1))) return rc;
    if ((rc = col_inv(h, Kd[4], h->Hd, 2, 0))) return rc;
    if ((rc = row_inv(h, h->Hd, X[3], 2, 0, sc))) return rc;
    const int parts = h->have_part && h->p.np > 0;
    if (parts) {
        particle_kv(h, Kp);
        if ((rc = col_inv(h, h->KV, h->Hd, 2, 0))) return rc;
        if ((rc = row_inv(h, h->Hd, X[4], 2, 0, sc))) return rc;
    }
    vk::DiagArgs da{h->p.alpha, h->p.omega, h->p.g, (double)1e-30f, h->N, parts};
    const int nb = 1024;
    hipLaunchKernelGGL(vk::k_diag_x, dim3(nb), dim3(256), 0, h->stream, (const cplx*)X[0], (const cplx*)X[1],
                       (const cplx*)X[2], (const cplx*)X[3], (const cplx*)X[4], X[1], X[2], X[5], X[6], da, h->dpart);
    hipLaunchKernelGGL(vk::k_reduce_fixed, dim3(1), dim3(256), 0, h->stream, (const double*)h->dpart, nb, 8, h->dsum);
    // spectra: FFT of w_x, w_y, q_x, q_y (kx >= 0 halves)
    if ((rc = fwd2d_half(h, X[1], h->Hd, Kd[0], 0, 0))) return rc;
    if ((rc = fwd2d_half(h, X[2], h->Hd, Kd[1], 0, 0))) return rc;
    if ((rc = fwd2d_half(h, X[5], h->Hd, Kd[2], 0, 0))) return rc;
    if ((rc = fwd2d_half(h, X[6], h->Hd, Kd[3], 0, 0))) return rc;
    const double en = 1.0 / ((double)h->N * (double)h->N);
    vk::SpecArgs sa{h->p.alpha, h->p.g, en * en, h->N, h->nshell};
    hipLaunchKernelGGL(vk::k_spectra, dim3(h->nshell + 1), dim3(256), 0, h->stream, (const int*)h->modes,
                       (const int*)h->offs, (const cplx*)Kd[0], (const cplx*)Kd[1], (const cplx*)Kd[2],
                       (const cplx*)Kd[3], (const cplx*)Kd[4], (const cplx*)h->K0, sa, h->dspec);
    if (parts) particle_forces(h, Kd[4], Kp, 1, 0);
    return 0;
}

// the diagnostics from the sums s[8], the spectra sp[6 (nshell + 1)] and
// the particle diagnostics dg[4] of diag_launch
void diag_final(const vk_gpe* h, const double* s, const double* sp, const double* dg, double* scal, double* spec,
                double* pscal)
{
    const double pi = 4.0 * std::atan(1.0);
    const int parts = h->have_part && h->p.np > 0;
    const double L = 2.0 * pi, dx = L / h->N, dvol = dx * dx, L2 = L * L, al = h->p.alpha;
    // final factors (src/diagnostics.f90:105-106,189-207,244)
    const double egrad = 2.0 * al * s[2] * dvol / L2;
    const double echem = 2.0 * al * s[3] * dvol / L2;
    const double eintr = 2.0 * al * s[4] * dvol / L2;
    if (scal) {
        scal[0] = egrad;
        scal[1] = echem;
        scal[2] = eintr;
        scal[3] = egrad + eintr;
        scal[4] = s[6] * dvol;
        scal[5] = 2.0 * al * s[0] * dvol / L2;
        scal[6] = 2.0 * al * s[1] * dvol / L2;
        scal[7] = sp[5 * (h->nshell + 1) + 0];
    }
    if (spec) std::memcpy(spec, sp, sizeof(double) * 6 * (h->nshell + 1));
    if (pscal) {
        for (int i = 0; i < 5; ++i) pscal[i] = 0.0;
        if (parts) {
            pscal[0] = 2.0 * al * s[5] * dvol / L2;
            pscal[1] = dg[0];
            pscal[2] = dg[1];
            pscal[3] = dg[2];
            pscal[4] = dg[3];
        }
    }
}

int diagnostics(vk_gpe* h, double* scal, double* spec, double* pscal)
{
    int rc = diag_launch(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    double s[8], dg[4] = {};
    HIPCHK(hipMemcpy(s, h->dsum, sizeof(s), hipMemcpyDeviceToHost));
    std::vector<double> sp(6 * (h->nshell + 1));
    HIPCHK(hipMemcpy(sp.data(), h->dspec, sizeof(double) * sp.size(), hipMemcpyDeviceToHost));
    if (h->have_part && h->p.np > 0)
        HIPCHK(hipMemcpy(dg, h->pst + 28 * h->p.np, sizeof(dg), hipMemcpyDeviceToHost));
    diag_final(h, s, sp.data(), dg, scal, spec, pscal);
    return 0;
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================

extern "C" {

const char* vk_last_error(void) { return g_err.c_str(); }
const char* vk_version(void) { return "vikshobha-mi355x 0.1 (gfx950)"; }

int vk_nshell(int N)
{
    const float a = 1.414f * (float)N;   // default-real literal (src/mod_grid.f90:77)
    return (int)((double)a / 2.0) + 1;
}

int vk_gpe_params_init(vk_gpe_params* p, int N, double dt, double xi_in, double csound, double rho_avg, double disp)
{
    if (!p || N < 32) return fail(VK_ERR_ARG, "params_init: bad arguments");
    std::memset(p, 0, sizeof(*p));
    const double pi = 4.0 * std::atan(1.0);
    const double L = 2.0 * pi;
    const double dx = L / N;
    p->N = N;
    p->dt = dt;
    p->xi = xi_in * dx;
    p->alpha = csound * p->xi * std::sqrt(rho_avg / 2.0);
    p->g = (csound * csound) / (2.0 * p->alpha);
    p->omega = p->g;
    p->disp = disp;
    p->mass = 1.0;
    p->force_dealias = 1;
    return 0;
}

int vk_gpe_create(vk_gpe** out, const vk_gpe_params* p)
{
    if (!out || !p) return fail(VK_ERR_ARG, "create: null argument");
    *out = nullptr;
    const int N = p->N;
    if (N < 32 || N > 4096 || (N & (N - 1))) return fail(VK_ERR_UNSUPPORTED, "N must be a power of two in [32, 4096]");
    if (p->np < 0 || p->np > 256) return fail(VK_ERR_ARG, "np must be in [0, 256]");
    if (p->np > 0 && !p->force_dealias)
        return fail(VK_ERR_UNSUPPORTED, "force_dealias=0 (no dealiasing of rho-hat in the particle force) is not implemented");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(VK_ERR_HIP, "no ROCm device available");
    vk_gpe* h = new vk_gpe();
    h->p = *p;
    h->N = N;
    h->nshell = vk_nshell(N);
    int TC = 0;
    dispatch(N, [&](auto n) { TC = CfgN<decltype(n)::value>::TC; return 0; });
    h->TC = TC;
    h->NP = N / TC;
    const int M = N / 3;
    h->geo.M = M;
    h->geo.Pp = (M + 1 + TC - 1) / TC;
    h->geo.Pn = (M + TC - 1) / TC;
    h->rs_part = use_real_space(*p);
    h->np_cap = p->np;
    if (h->geo.Pp + h->geo.Pn > h->NP) { h->geo.Pp = h->NP; h->geo.Pn = 0; }
    auto cleanup = [&](int rc) { vk_gpe_destroy(h); return rc; };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(VK_ERR_HIP, "hipStreamCreate failed"));
    const size_t fb = f2(h) * sizeof(cplx);
    cplx** bufs[] = {&h->K0, &h->K1, &h->K2, &h->Hpsi, &h->Hrho, &h->HN};
    for (auto b : bufs) {
        if (hipMalloc(b, fb) != hipSuccess || hipMemset(*b, 0, fb) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "hipMalloc failed (fields)"));
    }
    h->xpsi = N >= VK_XPSI_MIN_N;
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "device attribute query failed"));
        h->dev = dev;
        if (int rc = scratch_setup(dev)) return cleanup(rc);
        h->p1pf = p1_prefetch_ok(h);
        const char* t1 = std::getenv("VK_P1_TR1");
        const char* t3 = std::getenv("VK_P3_TR1");
        // Measured defaults: one row per workgroup in P1 and P3 at 2048^2
        // and 4096^2 (both with 4-column panels: 64-byte row pieces); 13%
        // faster per step than four rows at 2048^2, 3% than two at 4096^2.
        // VK_P1_TR1 / VK_P3_TR1 = 0 or 1 override; vk_gpe_autotune_rows
        // times every pair on the device at hand.
        h->tr1_p1 = t1 ? t1[0] == '1' : (N == 4096 || N == 2048);
        h->tr1_p3 = t3 ? t3[0] == '1' : (N == 4096 || N == 2048);
    }
    // psi rows: the single-field RK4/ARGLE path at N >= VK_XPSI_MIN_N and the
    // ensemble kernels (which also step single fields of N <= 1024)
    const bool rows = h->xpsi || (N >= 128 && N <= 1024);
    if (rows && hipMalloc(&h->Xs, fb) != hipSuccess) return cleanup(fail(VK_ERR_HIP, "hipMalloc failed (psi rows)"));
    if (hipMalloc(&h->tw, sizeof(cplx) * N) != hipSuccess) return cleanup(fail(VK_ERR_HIP, "hipMalloc tw"));
    // twiddles w_k = exp(-2 pi i k/N): long-double sin/cos on the first
    // octant, exact quadrant/octant symmetries elsewhere
    {
        std::vector<double2> tw(N);
        const long double pil = 3.141592653589793238462643383279502884L;
        for (int k = 0; k < N; ++k) {
            const int q = N / 4, quad = k / q;
            int a = k % q, flip = 0;
            if (2 * a > q) { a = q - a; flip = 1; }
            const long double t = 2.0L * pil * (long double)a / (long double)N;
            long double c = cosl(t), s = sinl(t);
            if (flip) { long double tmp = c; c = s; s = tmp; }
            for (int i = 0; i < quad; ++i) { long double tmp = c; c = -s; s = tmp; }
            tw[k] = make_double2((double)c, -(double)s);
        }
        if (hipMemcpy(h->tw, tw.data(), sizeof(double2) * N, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "twiddle upload"));
    }
    // residual partials: one pair per ARGLE column workgroup (at most one per column)
    if (hipMalloc(&h->part, sizeof(double) * 2 * N) != hipSuccess) return cleanup(fail(VK_ERR_HIP, "hipMalloc part"));
    if (hipMalloc(&h->dsum, sizeof(double) * 8) != hipSuccess) return cleanup(fail(VK_ERR_HIP, "hipMalloc dsum"));
    if (p->np > 0) {
        const int np = p->np;
        cplx** pb[] = {&h->HV, &h->KV, &h->Krho, &h->Vhat};
        for (auto b : pb) {
            if (hipMalloc(b, fb) != hipSuccess || hipMemset(*b, 0, fb) != hipSuccess)
                return cleanup(fail(VK_ERR_HIP, "hipMalloc failed (particles)"));
        }
        // force partials: spectral path 4 per row kx and particle; real-space
        // path 2 per row block, wave and particle
        size_t rows = 0, nw = 0;
        dispatch(N, [&](auto n) {
            using C = CfgN<decltype(n)::value>;
            rows = C::N;                      // one row per block at most
            nw = (C::ROWT + 63) / 64;
            return 0;
        });
        const size_t nfp = std::max((size_t)4 * np * N, 2 * rows * nw * np);
        if (hipMalloc(&h->EX, sizeof(cplx) * np * N) != hipSuccess ||
            hipMalloc(&h->EY, sizeof(cplx) * np * N) != hipSuccess ||
            hipMalloc(&h->fpart, sizeof(double) * nfp) != hipSuccess ||
            hipMalloc(&h->phit, rows * np) != hipSuccess ||
            hipMalloc(&h->pst, sizeof(double) * (30 * np + 8)) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "hipMalloc failed (particle tables)"));
        if (hipMemset(h->pst, 0, sizeof(double) * (30 * np + 8)) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "hipMemset failed (particle state)"));
        if (hipStreamCreateWithFlags(&h->pstream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_p3, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&h->ev_ps, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "particle stream/events"));
    } else {
        // P2/P3 take these pointers even when unused
        h->HV = h->Hrho;
        h->KV = h->K1;
        h->Krho = h->K1;
    }
    if (p->np == 0) {
        int rc = make_single_table(h);
        if (rc) return cleanup(rc);
    }
    if (hipDeviceSynchronize() != hipSuccess) return cleanup(fail(VK_ERR_HIP, "create: device sync"));
    *out = h;
    return 0;
}

void vk_gpe_destroy(vk_gpe* h)
{
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    cplx* fields[] = {h->K0, h->K1, h->K2, h->Hpsi, h->Hrho, h->HN, h->tw, h->Hd, h->Xs, h->Hdy, h->Xn, h->Kn};
    for (auto f : fields) if (f) hipFree(f);
    for (double* f : {h->Ux, h->Uy, h->omega_dev, h->npart}) if (f) hipFree(f);
    if (h->np_cap > 0) {
        cplx* pf[] = {h->HV, h->KV, h->Krho, h->Vhat, h->EX, h->EY};
        for (auto f : pf) if (f) hipFree(f);
        if (h->fpart) hipFree(h->fpart);
        if (h->phit) hipFree(h->phit);
        if (h->pst) hipFree(h->pst);
    }
    for (auto x : h->X) if (x) hipFree(x);
    for (auto k : h->Kd) if (k) hipFree(k);
    if (h->modes) hipFree(h->modes);
    if (h->offs) hipFree(h->offs);
    if (h->dspec) hipFree(h->dspec);
    if (h->dpart) hipFree(h->dpart);
    if (h->part) hipFree(h->part);
    if (h->dsum) hipFree(h->dsum);
    if (h->e1R) hipFree(h->e1R);
    for (int k = 0; k < 3; ++k) { if (h->e1map[k]) hipFree(h->e1map[k]); if (h->e1off[k]) hipFree(h->e1off[k]); }
    for (auto e : h->ev) hipEventDestroy(e);
    if (h->pstream) { hipStreamSynchronize(h->pstream); hipStreamDestroy(h->pstream); }
    if (h->ev_p3) hipEventDestroy(h->ev_p3);
    if (h->ev_ps) hipEventDestroy(h->ev_ps);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

static int sgle_sync_omega(vk_gpe* h);
static int ensure_hdy(vk_gpe* h);

static int need_x0(vk_gpe* h)
{
    if (!h->X[0]) HIPCHK(hipMalloc(&h->X[0], f2(h) * sizeof(cplx)));
    if (!h->Hd) {
        HIPCHK(hipMalloc(&h->Hd, f2(h) * sizeof(cplx)));
        HIPCHK(hipMemset(h->Hd, 0, f2(h) * sizeof(cplx)));
    }
    return 0;
}

static int set_from_x0(vk_gpe* h, int flags)
{
    int rc = fwd2d(h, h->X[0], h->Hd, h->K0, 0, 0);
    if (rc) return rc;
    h->hstate = 0;
    if (flags & VK_SYNC_DEALIAS) {
        if ((rc = ensure_rk4(h))) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int vk_gpe_set_psi(vk_gpe* h, const double* psi, int flags)
{
    if (!h || !psi) return fail(VK_ERR_ARG, "set_psi: null argument");
    int rc = need_x0(h);
    if (rc) return rc;
    // stream-ordered: a device-to-device hipMemcpy does not block the host and
    // is not ordered against this non-blocking stream
    HIPCHK(hipMemcpyAsync(h->X[0], psi, f2(h) * sizeof(cplx), hipMemcpyDefault, h->stream));
    return set_from_x0(h, flags);
}

int vk_gpe_set_kpsi(vk_gpe* h, const double* kpsi)
{
    if (!h || !kpsi) return fail(VK_ERR_ARG, "set_kpsi: null argument");
    HIPCHK(hipMemcpyAsync(h->K0, kpsi, f2(h) * sizeof(cplx), hipMemcpyDefault, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->hstate = 0;
    return 0;
}

int vk_gpe_get_psi(vk_gpe* h, double* psi, double* kpsi)
{
    if (!h) return fail(VK_ERR_ARG, "get_psi: null handle");
    int rc;
    if (psi) {
        if ((rc = need_x0(h))) return rc;
        if ((rc = ensure_any(h))) return rc;
        if ((rc = row_inv(h, h->Hpsi, h->X[0], h->hstate == 2 ? 1 : 0, 0, scale_of(h)))) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(psi, h->X[0], f2(h) * sizeof(cplx), hipMemcpyDefault, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    if (kpsi) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpyAsync(kpsi, h->K0, f2(h) * sizeof(cplx), hipMemcpyDefault, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

int vk_gpe_rk4_steps_async(vk_gpe* h, int nsteps)
{
    if (!h || nsteps < 0) return fail(VK_ERR_ARG, "rk4_steps: bad arguments");
    int rc = prep_rk4(h);
    if (rc) return rc;
    for (int i = 0; i < nsteps; ++i)
        if ((rc = rk4_step(h))) return rc;
    return 0;
}

int vk_gpe_rk4_steps(vk_gpe* h, int nsteps)
{
    int rc = vk_gpe_rk4_steps_async(h, nsteps);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int vk_gpe_argle_steps(vk_gpe* h, int nsteps, double* residual)
{
    if (!h || nsteps < 0) return fail(VK_ERR_ARG, "argle_steps: bad arguments");
    int rc = ensure_full(h);
    if (rc) return rc;
    for (int i = 0; i < nsteps; ++i)
        if ((rc = argle_step(h, residual && i == nsteps - 1))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    if ((rc = sgle_sync_omega(h))) return rc;
    if (residual && nsteps > 0) {
        double r[2];
        HIPCHK(hipMemcpy(r, h->dsum, sizeof(r), hipMemcpyDeviceToHost));
        *residual = std::sqrt(r[0]) / (h->p.dt * std::sqrt(r[1]));
    }
    return 0;
}

int vk_gpe_argle_until(vk_gpe* h, double tol, int max_steps, int check_every, int* steps_out, double* residual)
{
    if (!h || max_steps < 0 || check_every <= 0 || tol <= 0.0) return fail(VK_ERR_ARG, "argle_until: bad arguments");
    int rc = ensure_full(h);
    if (rc) return rc;
    int done = 0;
    double res = INFINITY;
    while (done < max_steps) {
        const int n = std::min(check_every, max_steps - done);
        for (int i = 0; i < n; ++i)
            if ((rc = argle_step(h, i == n - 1))) return rc;
        done += n;
        double r[2];
        HIPCHK(hipMemcpyAsync(r, h->dsum, sizeof(r), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        res = std::sqrt(r[0]) / (h->p.dt * std::sqrt(r[1]));
        if (!(res > tol)) break;   // converged (or NaN: stop and report)
    }
    HIPCHK(hipGetLastError());
    if ((rc = sgle_sync_omega(h))) return rc;
    if (steps_out) *steps_out = done;
    if (residual) *residual = res;
    return 0;
}

int vk_gpe_diagnostics(vk_gpe* h, double* scal, double* spec, double* pscal)
{
    if (!h) return fail(VK_ERR_ARG, "diagnostics: null handle");
    return diagnostics(h, scal, spec, pscal);
}

// particle_potential (src/gpe_particle.f90:3-90): V of one particle at the
// origin on the grid, forward transform, dealias -> Vhat
static int particle_potential(vk_gpe* h)
{
    const double pi = 4.0 * std::atan(1.0);
    const double L = 2.0 * pi, dx = L / h->N, c1 = 2 * pi / (2.0 * L);
    hipLaunchKernelGGL(vk::k_pot_x, dim3(1024), dim3(256), 0, h->stream, h->X[0], h->N, dx, c1, h->p.V0, h->p.d);
    int rc = fwd2d(h, h->X[0], h->Hd, h->Vhat, 0, 1);
    if (rc) return rc;
    h->have_part = 1;
    return 0;
}

int vk_gpe_particles_set(vk_gpe* h, const double* pos, const double* vel, const double* fext)
{
    if (!h) return fail(VK_ERR_ARG, "particles_set: null handle");
    const int np = h->p.np;
    if (np <= 0) return fail(VK_ERR_STATE, "particles_set: handle created with np = 0");
    int rc = need_x0(h);
    if (rc) return rc;
    // steps queued by vk_gpe_rk4_steps_async read and write the particle
    // state (the side-stream updates are joined into h->stream at the end of
    // every step): the copies go on h->stream, after them
    if (pos) HIPCHK(hipMemcpyAsync(h->pst, pos, sizeof(double) * 2 * np, hipMemcpyDefault, h->stream));
    if (vel) HIPCHK(hipMemcpyAsync(h->pst + 2 * np, vel, sizeof(double) * 2 * np, hipMemcpyDefault, h->stream));
    if (fext) HIPCHK(hipMemcpyAsync(h->pst + 4 * np, fext, sizeof(double) * 2 * np, hipMemcpyDefault, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!h->have_part && (rc = particle_potential(h))) return rc;
    h->kv_valid = 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

}  // extern "C"

// accessors for the library's other translation units (not part of the C ABI)
cplx* vk_internal_kpsi(vk_gpe* h) { return h ? h->K0 : nullptr; }
int vk_internal_N(const vk_gpe* h) { return h ? h->N : 0; }
hipStream_t vk_internal_stream(const vk_gpe* h) { return h ? h->stream : nullptr; }

extern "C" {

int vk_gpe_set_advection(vk_gpe* h, int taylor_green, double vnx, double vny)
{
    if (!h || (taylor_green != 0 && taylor_green != 1) || !std::isfinite(vnx) || !std::isfinite(vny))
        return fail(VK_ERR_ARG, "set_advection: taylor_green must be 0 or 1 and the velocity finite");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (taylor_green || vnx != 0.0 || vny != 0.0) {
        int rc = ensure_hdy(h);
        if (rc) return rc;
    }
    h->adv_tg = taylor_green;
    h->adv_vx = vnx;
    h->adv_vy = vny;
    return 0;
}

static int ensure_hdy(vk_gpe* h)
{
    if (!h->Hdy) {
        HIPCHK(hipMalloc(&h->Hdy, f2(h) * sizeof(cplx)));
        HIPCHK(hipMemset(h->Hdy, 0, f2(h) * sizeof(cplx)));
    }
    return 0;
}

static int ensure_field_bufs(vk_gpe* h)
{
    if (!h->Ux) HIPCHK(hipMalloc(&h->Ux, f2(h) * sizeof(double)));
    if (!h->Uy) HIPCHK(hipMalloc(&h->Uy, f2(h) * sizeof(double)));
    return ensure_hdy(h);
}

int vk_gpe_set_advection_field(vk_gpe* h, const double* ux, const double* uy)
{
    if (!h || (!ux) != (!uy)) return fail(VK_ERR_ARG, "set_advection_field: both components or neither");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!ux) {
        h->adv_field = 0;
        return 0;
    }
    int rc = ensure_field_bufs(h);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->Ux, ux, f2(h) * sizeof(double), hipMemcpyDefault, h->stream));
    HIPCHK(hipMemcpyAsync(h->Uy, uy, f2(h) * sizeof(double), hipMemcpyDefault, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->adv_field = 1;
    return 0;
}

int vk_gpe_set_vortex_advection(vk_gpe* h, int nv, const double* xs, const double* ys, const int* qs, int nimg)
{
    if (!h || nv <= 0 || !xs || !ys || !qs || nimg < 0) return fail(VK_ERR_ARG, "set_vortex_advection: bad arguments");
    for (int j = 0; j < nv; ++j)
        if (qs[j] != 1 && qs[j] != -1) return fail(VK_ERR_ARG, "set_vortex_advection: charges must be +-1");
    int rc = ensure_field_bufs(h);
    if (rc) return rc;
    double* d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(double) * 2 * nv + sizeof(int) * nv));
    HIPCHK(hipMemcpy(d, xs, sizeof(double) * nv, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + nv, ys, sizeof(double) * nv, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d + 2 * nv, qs, sizeof(int) * nv, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(vk::k_vortex_velocity, dim3(1024), dim3(256), 0, h->stream, h->Ux, h->Uy, h->N, nv,
                       (const double*)d, (const double*)(d + nv), (const int*)(d + 2 * nv), h->p.xi, nimg,
                       h->p.alpha);
    const hipError_t e = hipStreamSynchronize(h->stream);
    hipFree(d);
    HIPCHK(e);
    HIPCHK(hipGetLastError());
    h->adv_field = 1;
    return 0;
}

int vk_vortex_velocity_host(int N, int nv, const double* xs, const double* ys, const int* qs, double xi,
                            double alpha, int nimg, double* ux, double* uy)
{
    if (N <= 0 || nv <= 0 || !xs || !ys || !qs || !ux || !uy || nimg < 0)
        return fail(VK_ERR_ARG, "vortex_velocity_host: bad arguments");
    const double dx = 2.0 * 3.14159265358979323846 / N;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i)
            vk::vortex_velocity_point(i * dx, j * dx, nv, xs, ys, qs, xi, nimg, alpha, ux + (size_t)j * N + i,
                                      uy + (size_t)j * N + i);
    return 0;
}

int vk_gpe_set_sgle(vk_gpe* h, int on, double oneybeta, double nuN, double chempot0, double ktrunc,
                    unsigned long long seed)
{
    if (!h || (on != 0 && on != 1) || !(oneybeta >= 0.0) || !std::isfinite(nuN) || !std::isfinite(chempot0) ||
        !std::isfinite(ktrunc) || !std::isfinite(oneybeta))
        return fail(VK_ERR_ARG, "set_sgle: on must be 0 or 1, oneybeta >= 0 and the parameters finite");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (!on) {
        h->sgle = 0;
        return 0;
    }
    int rc = need_x0(h);
    if (rc) return rc;
    if (!h->Xn) HIPCHK(hipMalloc(&h->Xn, f2(h) * sizeof(cplx)));
    if (!h->Kn) HIPCHK(hipMalloc(&h->Kn, f2(h) * sizeof(cplx)));
    if (!h->omega_dev) HIPCHK(hipMalloc(&h->omega_dev, sizeof(double)));
    if (!h->npart) HIPCHK(hipMalloc(&h->npart, sizeof(double) * kNormBlocks));
    HIPCHK(hipMemcpy(h->omega_dev, &chempot0, sizeof(double), hipMemcpyHostToDevice));
    h->p.omega = chempot0;
    h->sgle = 1;
    h->sg_beta = oneybeta;
    h->sg_nuN = nuN;
    h->sg_ktrunc = ktrunc;
    h->sg_seed = seed;
    h->sg_draw = 0;
    return 0;
}

// the chemical potential the SGLE steps have reached, into the handle's
// parameters (RK4 after SGLE runs at it, as the reference's omega_chem)
static int sgle_sync_omega(vk_gpe* h)
{
    if (!h->sgle) return 0;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(&h->p.omega, h->omega_dev, sizeof(double), hipMemcpyDeviceToHost));
    if (h->e1R) {
        const vk::EnsReal x = ens_entry(h);
        HIPCHK(hipMemcpy(h->e1R, &x, sizeof(x), hipMemcpyHostToDevice));
    }
    return 0;
}

int vk_gpe_get_chempot(vk_gpe* h, double* omega)
{
    if (!h || !omega) return fail(VK_ERR_ARG, "get_chempot: null argument");
    int rc = sgle_sync_omega(h);
    if (rc) return rc;
    *omega = h->p.omega;
    return 0;
}

int vk_gpe_argle_step_noise(vk_gpe* h, const double* noise_xy, double* residual)
{
    if (!h || !noise_xy) return fail(VK_ERR_ARG, "argle_step_noise: null argument");
    if (!h->sgle) return fail(VK_ERR_STATE, "argle_step_noise: SGLE is off (vk_gpe_set_sgle)");
    int rc = ensure_full(h);
    if (rc) return rc;
    if ((rc = argle_step(h, residual != nullptr, noise_xy))) return rc;
    if ((rc = sgle_sync_omega(h))) return rc;
    HIPCHK(hipGetLastError());
    if (residual) {
        double r[2];
        HIPCHK(hipMemcpy(r, h->dsum, sizeof(r), hipMemcpyDeviceToHost));
        *residual = std::sqrt(r[0]) / (h->p.dt * std::sqrt(r[1]));
    }
    return 0;
}

int vk_sgle_noise(vk_gpe* h, unsigned long long draw, double* out_xy)
{
    if (!h || !out_xy) return fail(VK_ERR_ARG, "sgle_noise: null argument");
    if (!h->sgle) return fail(VK_ERR_STATE, "sgle_noise: SGLE is off (vk_gpe_set_sgle)");
    hipLaunchKernelGGL(vk::k_sgle_noise, dim3(1024), dim3(256), 0, h->stream, h->Xn, h->N, draw, h->sg_seed, 1.0, 1.0,
                       1);
    HIPCHK(hipMemcpyAsync(out_xy, h->Xn, f2(h) * sizeof(cplx), hipMemcpyDefault, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int vk_gpe_set_dt(vk_gpe* h, double dt)
{
    if (!h || !(dt > 0.0) || !std::isfinite(dt)) return fail(VK_ERR_ARG, "set_dt: dt must be positive and finite");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->p.dt = dt;
    if (h->e1R) {
        const vk::EnsReal x = ens_entry(h);
        HIPCHK(hipMemcpy(h->e1R, &x, sizeof(x), hipMemcpyHostToDevice));
    }
    return 0;
}

int vk_gpe_set_obstacles(vk_gpe* h, int np, const double* pos, double V0, double d)
{
    if (!h) return fail(VK_ERR_ARG, "set_obstacles: null handle");
    if (np < 0 || np > h->np_cap)
        return fail(VK_ERR_ARG, "set_obstacles: np must be in [0, np of the params the handle was created with]");
    if (np > 0 && (!pos || !(d > 0.0) || !std::isfinite(V0)))
        return fail(VK_ERR_ARG, "set_obstacles: positions, d > 0 and a finite V0 are required");
    int rc = need_x0(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->p.np = np;
    h->p.V0 = V0;
    h->p.d = d;
    h->rs_part = use_real_space(h->p);
    h->have_part = 0;
    h->kv_valid = 0;
    if (np == 0) return 0;
    std::vector<double> st(30 * (size_t)np + 8, 0.0);
    std::copy(pos, pos + 2 * np, st.begin());
    HIPCHK(hipMemcpyAsync(h->pst, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice, h->stream));
    if ((rc = particle_potential(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return 0;
}

int vk_gpe_particles_get(vk_gpe* h, double* pos, double* vel, double* frc)
{
    if (!h) return fail(VK_ERR_ARG, "particles_get: null handle");
    const int np = h->p.np;
    if (np <= 0) return fail(VK_ERR_STATE, "particles_get: handle created with np = 0");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (pos) HIPCHK(hipMemcpy(pos, h->pst, sizeof(double) * 2 * np, hipMemcpyDeviceToHost));
    if (vel) HIPCHK(hipMemcpy(vel, h->pst + 2 * np, sizeof(double) * 2 * np, hipMemcpyDeviceToHost));
    if (frc) HIPCHK(hipMemcpy(frc, h->pst + 26 * np, sizeof(double) * 2 * np, hipMemcpyDeviceToHost));
    return 0;
}

int vk_fft2d(vk_gpe* h, int dir, const double* in, double* out)
{
    if (!h || !in || !out || (dir != -1 && dir != 1)) return fail(VK_ERR_ARG, "fft2d: bad arguments");
    int rc = need_x0(h);
    if (rc) return rc;
    if (!h->Kd[0]) HIPCHK(hipMalloc(&h->Kd[0], f2(h) * sizeof(cplx)));
    const size_t fb = f2(h) * sizeof(cplx);
    if (dir < 0) {
        HIPCHK(hipMemcpy(h->X[0], in, fb, hipMemcpyHostToDevice));
        if ((rc = fwd2d(h, h->X[0], h->Hd, h->Kd[0], 0, 0))) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, h->Kd[0], fb, hipMemcpyDeviceToHost));
    } else {
        HIPCHK(hipMemcpy(h->Kd[0], in, fb, hipMemcpyHostToDevice));
        if ((rc = col_inv(h, h->Kd[0], h->Hd, 0, 0))) return rc;
        if ((rc = row_inv(h, h->Hd, h->X[0], 0, 0, scale_of(h)))) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, h->X[0], fb, hipMemcpyDeviceToHost));
    }
    return 0;
}

// The transforms of vk_fft2d between caller-owned device buffers, with the
// half-spectral intermediate in the static scratch instead of library
// memory: no allocation of any kind.  Stream-ordered on the handle's stream
// (returns once the launches are queued).
int vk_fft2d_noalloc(vk_gpe* h, int dir, const double* in, double* out)
{
    if (!h || !in || !out || (dir != -1 && dir != 1) || (const void*)in == (const void*)out)
        return fail(VK_ERR_ARG, "fft2d_noalloc: bad arguments (distinct device buffers, dir = -1 or +1)");
    if (h->N > VK_NOALLOC_MAX_N)
        return fail(VK_ERR_UNSUPPORTED, "fft2d_noalloc: N = " + std::to_string(h->N) + " exceeds the " +
                                            std::to_string(VK_NOALLOC_MAX_N) + "^2 scratch");
    for (const void* q : {(const void*)in, (const void*)out}) {
        hipPointerAttribute_t a{};
        if (hipPointerGetAttributes(&a, q) != hipSuccess || a.type != hipMemoryTypeDevice) {
            (void)hipGetLastError();
            return fail(VK_ERR_ARG, "fft2d_noalloc: in and out must be device memory");
        }
    }
    cplx* H = g_scratch_ptr[h->dev];
    if (!H) return fail(VK_ERR_STATE, "fft2d_noalloc: scratch not set up on this device");
    const cplx* ci = reinterpret_cast<const cplx*>(in);
    cplx* co = reinterpret_cast<cplx*>(out);
    int rc;
    if (dir < 0) rc = fwd2d(h, ci, H, co, 0, 0);
    else if (!(rc = col_inv(h, ci, H, 0, 0))) rc = row_inv(h, H, co, 0, 0, scale_of(h));
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    return 0;
}

int vk_ic_vortices_host(int N, int nv, const double* xs, const double* ys, const int* qs, double xi, int nimg,
                        double* psi)
{
    if (N < 2 || nv < 0 || !psi || (nv > 0 && (!xs || !ys || !qs))) return fail(VK_ERR_ARG, "ic: bad arguments");
    const double dx = 2.0 * 3.14159265358979323846 / N;
    for (int iy = 0; iy < N; ++iy)
        for (int ix = 0; ix < N; ++ix) vk::ic_point(ix * dx, iy * dx, nv, xs, ys, qs, xi, nimg, psi + 2 * ((size_t)iy * N + ix));
    return 0;
}

int vk_gpe_set_ic_vortices(vk_gpe* h, int nv, const double* xs, const double* ys, const int* qs, int nimg, int flags)
{
    if (!h || nv < 0 || (nv > 0 && (!xs || !ys || !qs))) return fail(VK_ERR_ARG, "set_ic_vortices: bad arguments");
    int rc = need_x0(h);
    if (rc) return rc;
    double *dx = nullptr, *dy = nullptr;
    int* dq = nullptr;
    const int n = nv > 0 ? nv : 1;
    HIPCHK(hipMalloc(&dx, sizeof(double) * n));
    HIPCHK(hipMalloc(&dy, sizeof(double) * n));
    HIPCHK(hipMalloc(&dq, sizeof(int) * n));
    if (nv > 0) {
        HIPCHK(hipMemcpy(dx, xs, sizeof(double) * nv, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dy, ys, sizeof(double) * nv, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dq, qs, sizeof(int) * nv, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(vk::k_ic, dim3(2048), dim3(256), 0, h->stream, h->X[0], h->N, nv, (const double*)dx,
                       (const double*)dy, (const int*)dq, h->p.xi, nimg);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    hipFree(dx);
    hipFree(dy);
    hipFree(dq);
    return set_from_x0(h, flags);
}

void* vk_gpe_stream(vk_gpe* h) { return h ? (void*)h->stream : nullptr; }

int vk_pass_kinds(void) { return PK_COUNT; }
const char* vk_pass_name(int kind) { return (kind >= 0 && kind < PK_COUNT) ? kPassNames[kind] : "?"; }

int vk_gpe_pass_times(vk_gpe* h, int nsteps, double* ms_out, int nkinds)
{
    if (!h || nsteps <= 0 || !ms_out) return fail(VK_ERR_ARG, "pass_times: bad arguments");
    int rc = prep_rk4(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->timing = true;
    h->evn = 0;
    h->evkind.clear();
    for (int i = 0; i < nsteps && !rc; ++i) rc = rk4_step(h);
    h->timing = false;
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    std::vector<double> sum(PK_COUNT, 0.0);
    std::vector<int> cnt(PK_COUNT, 0);
    for (size_t i = 0; i < h->evkind.size(); ++i) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, h->ev[2 * i], h->ev[2 * i + 1]));
        sum[h->evkind[i]] += ms;
        cnt[h->evkind[i]]++;
    }
    for (int k = 0; k < nkinds && k < PK_COUNT; ++k) ms_out[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    return 0;
}

int vk_gpe_set_row_variants(vk_gpe* h, int p1, int p3)
{
    if (!h || p1 < 0 || p1 > 2 || p3 < 0 || p3 > 1) return fail(VK_ERR_ARG, "set_row_variants: bad arguments");
    if (!h->xpsi) return fail(VK_ERR_UNSUPPORTED, "set_row_variants: the row-pass variants exist for the psi-row path only");
    if (p1 == 2 && !p1_prefetch_fits(h))
        return fail(VK_ERR_UNSUPPORTED, "set_row_variants: the prefetching P1 does not fit this size or device");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->p1pf = p1 == 2;
    h->tr1_p1 = p1 == 1;
    h->tr1_p3 = p3;
    return 0;
}

int vk_gpe_get_row_variants(vk_gpe* h, int* p1, int* p3)
{
    if (!h) return fail(VK_ERR_ARG, "get_row_variants: null");
    if (p1) *p1 = h->p1pf ? 2 : (h->tr1_p1 ? 1 : 0);
    if (p3) *p3 = h->tr1_p3;
    return 0;
}

int vk_gpe_autotune_rows(vk_gpe* h, int nsteps, int* choice, double* ms_out)
{
    if (!h || nsteps <= 0) return fail(VK_ERR_ARG, "autotune_rows: bad arguments");
    if (!h->xpsi) return fail(VK_ERR_UNSUPPORTED, "autotune_rows: the row-pass variants exist for the psi-row path only");
    int rc = ensure_rk4(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    // save the state: kpsi, its y-inverse, the particle state
    const size_t nf = f2(h), npst = h->pst ? (size_t)30 * h->np_cap + 8 : 0;
    cplx* sv = nullptr;
    double* sp = nullptr;
    HIPCHK(hipMalloc(&sv, 2 * nf * sizeof(cplx)));
    if (npst && hipMalloc(&sp, npst * sizeof(double)) != hipSuccess) {
        hipFree(sv);
        return fail(VK_ERR_HIP, "autotune_rows: hipMalloc");
    }
    auto release = [&]() { hipFree(sv); if (sp) hipFree(sp); };
    auto copy = [&](bool save) -> int {
        HIPCHK(hipMemcpyAsync(save ? sv : h->K0, save ? h->K0 : sv, nf * sizeof(cplx), hipMemcpyDeviceToDevice,
                              h->stream));
        HIPCHK(hipMemcpyAsync(save ? sv + nf : h->Hpsi, save ? h->Hpsi : sv + nf, nf * sizeof(cplx),
                              hipMemcpyDeviceToDevice, h->stream));
        if (npst)
            HIPCHK(hipMemcpyAsync(save ? sp : h->pst, save ? h->pst : sp, npst * sizeof(double),
                                  hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    };
    if ((rc = copy(true))) { release(); return rc; }
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        release();
        return fail(VK_ERR_HIP, "autotune_rows: events");
    }
    const int keep1 = h->p1pf ? 2 : (h->tr1_p1 ? 1 : 0), keep3 = h->tr1_p3;
    double best = 1e300;
    int b1 = keep1, b3 = keep3;
    for (int p1 = 0; p1 <= 2 && !rc; ++p1)
        for (int p3 = 0; p3 <= 1 && !rc; ++p3) {
            if (ms_out) ms_out[2 * p1 + p3] = -1.0;
            if (p1 == 2 && !p1_prefetch_fits(h)) continue;
            h->p1pf = p1 == 2;
            h->tr1_p1 = p1 == 1;
            h->tr1_p3 = p3;
            if ((rc = vk_gpe_rk4_steps_async(h, 1))) break;   // warm
            if (hipEventRecord(a, h->stream) != hipSuccess) { rc = fail(VK_ERR_HIP, "autotune_rows: event"); break; }
            if ((rc = vk_gpe_rk4_steps_async(h, nsteps))) break;
            if (hipEventRecord(b, h->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess) {
                rc = fail(VK_ERR_HIP, "autotune_rows: event");
                break;
            }
            float ms = 0.f;
            hipEventElapsedTime(&ms, a, b);
            const double per = ms / nsteps;
            if (ms_out) ms_out[2 * p1 + p3] = per;
            if (per < best) { best = per; b1 = p1; b3 = p3; }
        }
    hipEventDestroy(a);
    hipEventDestroy(b);
    h->p1pf = b1 == 2;
    h->tr1_p1 = b1 == 1;
    h->tr1_p3 = b3;
    // the starting state back, also after a failed trial (whose error code
    // is the one returned)
    hipStreamSynchronize(h->stream);
    const int rc2 = copy(false);
    if (!rc) rc = rc2;
    release();
    if (rc) return rc;
    h->hstate = 2;
    h->kv_valid = 0;
    if (choice) { choice[0] = b1; choice[1] = b3; }
    return 0;
}

double vk_gpe_pass_bytes(vk_gpe* h, int kind)
{
    if (!h) return 0.0;
    const double F = 16.0 * (double)h->N * (double)h->N;
    const double fp = (double)(h->geo.Pp * h->TC) / h->N;                        // kx >= 0 half columns
    const double fr = (double)((h->geo.Pp + h->geo.Pn) * h->TC) / h->N;         // retained columns
    // retained disk modes
    long long disk = 0;
    for (int kx = -h->N / 2; kx < h->N / 2; ++kx)
        for (int ky = -h->N / 2; ky < h->N / 2; ++ky)
            if (!masked_host(kx, ky, h->N)) ++disk;
    const double D = 16.0 * (double)disk;
    // the RK4 step of a single field of N <= 1024 without particles runs the
    // ensemble kernels, which pass psi rows from P1 to P3 as well
    const bool rows = h->xpsi || (h->e1R && !(h->have_part && h->p.np > 0));
    switch (kind) {
    case PK_P1: return F * (fr + fp + (rows ? 1.0 : 0.0));   // + psi rows written for P3
    case PK_P2: return F * 2.0 * fp;
    case PK_P3: return F * (fp + (rows ? 1.0 : fr) + fr);    // psi rows instead of Hpsi
    // K accesses per mode, stages 1-4: K0 read + K2 write; K0, K2 read + K1,
    // K2 write; K0, K2 read + K2 write (str2 is not read in stage 3); K0, K1,
    // K2 read + K0 write
    case PK_P4: return F * (fr + fr) + D * (2 + 4 + 3 + 4) / 4.0;
    default: return 0.0;
    }
}

}  // extern "C"

// =============================================================================
// Ensemble of independent realizations (batched launches, graph replay)
// =============================================================================

struct vk_ens {
    std::vector<vk_gpe*> h;
    hipStream_t stream = nullptr;
    vk::EnsReal* dR = nullptr;
    // block maps: 0 row passes (P1, P3), 1 P2, 2 P4; per-size maps for mode 1
    int* dmap[3] = {};
    int* doff[3] = {};
    int nblk[3] = {};
    std::vector<int> sizes;                       // distinct N
    std::map<int, std::vector<int>> size_reals;   // N -> realization ids
    std::map<int, std::array<int*, 6>> smaps;     // N -> {map0, off0, map1, off1, map2, off2}
    std::map<int, std::array<int, 3>> snblk;
    size_t lds[4] = {};                           // max dynamic LDS per pass
    // 0: one launch per pass for all sizes; 1: one per size; 2: per realization;
    // 3: one persistent launch per step (all passes, grid barriers between)
    int mode = 0;
    int use_graph = 1;
    hipGraphExec_t gexec[4] = {};
    unsigned* dbar = nullptr;   // the persistent step's barrier counters
    int* derr = nullptr;        // set by a barrier wait that ran out of time
    int pgrid = 0;              // persistent grid: resident workgroups, multiple of 8
    size_t plds = 0;
    long long plimit = 0;       // barrier wait limit in wall-clock ticks
    double* hstage = nullptr;   // pinned staging of every member's diagnostics
    double* dstage = nullptr;   // the same on the device
    size_t nstage = 0;
    void* dtab = nullptr;       // operand tables of the batched diagnostics
    void* htab = nullptr;       // their pinned host staging
    size_t ntab = 0;
    int synced = 0;
    std::map<int, int> var;                       // N -> shape variant
};

namespace {

// panel geometry of panel width TC (as vk_gpe_create computes it)
Geo geo_of(int N, int TC)
{
    Geo g{};
    const int NP = N / TC, M = N / 3;
    g.M = M;
    g.Pp = (M + 1 + TC - 1) / TC;
    g.Pn = (M + TC - 1) / TC;
    if (g.Pp + g.Pn > NP) { g.Pp = NP; g.Pn = 0; }
    return g;
}

// call f with the shape type of (N, var)
template <class F>
int dispatch_var(int N, int var, F&& f)
{
    switch (N * 2 + var) {
    case 256: return f(vk::CfgE<128, 0>{});
    case 512: return f(vk::CfgE<256, 0>{});
    case 513: return f(vk::CfgE<256, 1>{});
    case 1024: return f(vk::CfgE<512, 0>{});
    case 1025: return f(vk::CfgE<512, 1>{});
    case 2048: return f(vk::CfgE<1024, 0>{});
    case 2049: return f(vk::CfgE<1024, 1>{});
    default: return fail(VK_ERR_UNSUPPORTED, "no such ensemble shape");
    }
}

template <class C> void ens_lds(size_t* l)
{
    l[0] = std::max(l[0], (size_t)C::TR * C::ROWRS * sizeof(double));
    l[1] = std::max(l[1], (size_t)C::TC * C::COLRS * sizeof(double));
    l[2] = std::max(l[2], (size_t)C::TR * C::ROWRS * sizeof(double) + 16);   // f in registers (psi rows)
    l[3] = std::max(l[3], (size_t)C::TC * C::COLRS * sizeof(double));
}

// Block maps for the realizations in ids: for pass kinds 0 (row passes),
// 1 (P2) and 2 (P4), map[b] = realization of block b (runs contiguous, in
// ids order) and off[2r], off[2r+1] = [start, end) of r's run.
int build_maps(const std::vector<vk_gpe*>& hs, const std::vector<int>& ids, int** dmap, int** doff, int* nblk)
{
    for (int k = 0; k < 3; ++k) {
        std::vector<int> map;
        std::vector<int> se(2 * hs.size(), 0);
        for (int r : ids) {
            const vk_gpe* h = hs[r];
            int TR = 0;
            dispatch_var(h->N, h->ens_var, [&](auto c) { TR = decltype(c)::TR; return 0; });
            const Geo& gv = h->ens_var ? h->geo_v : h->geo;
            const int nb = k == 0 ? h->N / TR : (k == 1 ? gv.Pp : gv.Pp + gv.Pn);
            se[2 * r] = (int)map.size();
            for (int b = 0; b < nb; ++b) map.push_back(r);
            se[2 * r + 1] = (int)map.size();
        }
        nblk[k] = (int)map.size();
        HIPCHK(hipMalloc(&dmap[k], sizeof(int) * std::max<size_t>(1, map.size())));
        HIPCHK(hipMalloc(&doff[k], sizeof(int) * se.size()));
        HIPCHK(hipMemcpy(dmap[k], map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(doff[k], se.data(), sizeof(int) * se.size(), hipMemcpyHostToDevice));
    }
    return 0;
}

// set the LDS limit of the single-field pass kernels of size C before any
// graph capture
template <class C> void preconfig_passes()
{
    set_lds(vk::k_p1<C>, (size_t)C::TR * C::ROWRS * sizeof(double));
    set_lds(vk::k_p2<C>, (size_t)C::TC * C::COLRS * sizeof(double));
    set_lds(vk::k_p3<C, 0>, (size_t)C::TR * C::ROWRS * sizeof(double) + (size_t)C::S * C::ROWT * sizeof(double) + 16);
    for (auto k : {vk::k_p4<C, 1>, vk::k_p4<C, 2>, vk::k_p4<C, 3>, vk::k_p4<C, 4>})
        set_lds(k, (size_t)C::TC * C::COLRS * sizeof(double));
}

vk::EnsReal ens_entry(const vk_gpe* h)
{
    vk::EnsReal x{};
    x.N = h->N;
    x.var = h->ens_var;
    x.g = h->ens_var ? h->geo_v : h->geo;
    x.K0 = h->K0; x.K1 = h->K1; x.K2 = h->K2; x.Hpsi = h->Hpsi; x.Hrho = h->Hrho; x.HN = h->HN; x.Xs = h->Xs;
    x.tw = h->tw;
    x.alpha = h->p.alpha;
    x.omega = h->p.omega;
    x.sc = scale_of(h);
    x.gsc = h->p.g * x.sc;
    x.dt = h->p.dt;
    x.pref = make_double2(-h->p.disp, -1.0);
    return x;
}

int make_single_table(vk_gpe* h)
{
    if (h->N < 128 || h->N > 1024) return 0;
    // the table and block maps of the single-field shape svar (build_maps
    // and ens_entry read the member shape fields, borrowed here)
    const int keep_var = h->ens_var;
    const Geo keep_geo = h->geo_v;
    h->ens_var = h->svar;
    h->geo_v = h->geo_s;
    const vk::EnsReal x = ens_entry(h);
    for (int k = 0; k < 3; ++k) {
        if (h->e1map[k]) { hipFree(h->e1map[k]); h->e1map[k] = nullptr; }
        if (h->e1off[k]) { hipFree(h->e1off[k]); h->e1off[k] = nullptr; }
    }
    std::vector<vk_gpe*> hs{h};
    int rc = 0;
    if (!h->e1R && hipMalloc(&h->e1R, sizeof(vk::EnsReal)) != hipSuccess) rc = fail(VK_ERR_HIP, "hipMalloc e1R");
    if (!rc && hipMemcpy(h->e1R, &x, sizeof(x), hipMemcpyHostToDevice) != hipSuccess) rc = fail(VK_ERR_HIP, "e1R copy");
    if (!rc) rc = build_maps(hs, std::vector<int>{0}, h->e1map, h->e1off, h->e1nblk);
    h->ens_var = keep_var;
    h->geo_v = keep_geo;
    if (rc) return rc;
    for (auto& l : h->e1lds) l = 0;
    dispatch_var(h->N, h->svar, [&](auto c) { ens_lds<decltype(c)>(h->e1lds); return 0; });
    set_lds(vk::k_ens_p1, h->e1lds[0]);
    set_lds(vk::k_ens_p2, h->e1lds[1]);
    set_lds(vk::k_ens_p3, h->e1lds[2]);
    set_lds(vk::k_ens_p4, h->e1lds[3]);
    for (auto k : {vk::k_ens_p4s<1>, vk::k_ens_p4s<2>, vk::k_ens_p4s<3>, vk::k_ens_p4s<4>}) set_lds(k, h->e1lds[3]);
    return 0;
}

// Before RK4 steps: K0 dealiased and Hpsi = IFFT_y(K0) on the retained
// panels in the layout the steps use: the single-field shape's for the
// one-member ensemble path (hstate 3 with svar = 1), else Cfg<N>'s
// (hstate 2).
int prep_rk4(vk_gpe* h)
{
    const bool e1 = h->e1R && !(h->have_part && h->p.np > 0);
    if (!(e1 && h->svar)) return ensure_rk4(h);
    if (h->hstate == 3) return 0;
    int rc = ensure_rk4(h);
    if (rc) return rc;
    rc = dispatch_var(h->N, h->svar, [&](auto c) {
        using C = decltype(c);
        const size_t lds = (size_t)C::TC * C::COLRS * sizeof(double);
        set_lds(vk::k_col_inv<C>, lds);
        hipLaunchKernelGGL(vk::k_col_inv<C>, dim3(h->geo_s.Pp + h->geo_s.Pn), dim3(C::COLT), lds, h->stream,
                           (const cplx*)h->K0, h->Hpsi, (const cplx*)h->tw, h->geo_s, 1, 0);
        return 0;
    });
    if (rc) return rc;
    h->hstate = 3;
    return 0;
}

// the ensemble P4 of one stage: a kernel per stage, or one kernel taking
// the stage at run time (VK_ENS_P4_RUNTIME_STAGE=1)
#ifndef VK_ENS_P4_RUNTIME_STAGE
#define VK_ENS_P4_RUNTIME_STAGE 0
#endif
void ens_p4_launch(hipStream_t st, int nb, size_t lds, const vk::EnsReal* R, const int* m, const int* o, int stage)
{
    if (VK_ENS_P4_RUNTIME_STAGE) {
        hipLaunchKernelGGL(vk::k_ens_p4, dim3(nb), dim3(256), lds, st, R, m, o, stage);
        return;
    }
    auto k = stage == 1 ? vk::k_ens_p4s<1> : stage == 2 ? vk::k_ens_p4s<2> : stage == 3 ? vk::k_ens_p4s<3>
                                                                                   : vk::k_ens_p4s<4>;
    hipLaunchKernelGGL(k, dim3(nb), dim3(256), lds, st, R, m, o);
}

// launch the four passes of every stage over the given block maps
void ens_passes(hipStream_t st, const vk::EnsReal* R, int* const* m, int* const* o, const int* nb, const size_t* lds,
                vk_gpe* th)
{
    for (int stage = 1; stage <= 4; ++stage) {
        if (th) ev_begin(th);
        hipLaunchKernelGGL(vk::k_ens_p1, dim3(nb[0]), dim3(256), lds[0], st, R, (const int*)m[0], (const int*)o[0]);
        if (th) ev_end(th, PK_P1), ev_begin(th);
        hipLaunchKernelGGL(vk::k_ens_p2, dim3(nb[1]), dim3(256), lds[1], st, R, (const int*)m[1], (const int*)o[1]);
        if (th) ev_end(th, PK_P2), ev_begin(th);
        hipLaunchKernelGGL(vk::k_ens_p3, dim3(nb[0]), dim3(256), lds[2], st, R, (const int*)m[0], (const int*)o[0]);
        if (th) ev_end(th, PK_P3), ev_begin(th);
        ens_p4_launch(st, nb[2], lds[3], R, (const int*)m[2], (const int*)o[2], stage);
        if (th) ev_end(th, PK_P4);
    }
}

int ens1_step(vk_gpe* h)
{
    ens_passes(h->stream, h->e1R, h->e1map, h->e1off, h->e1nblk, h->e1lds, h);
    h->kv_valid = 0;
    return 0;
}

}  // namespace

namespace {

// shape variant per (N, GPU model) chosen by vk_ens_autotune_shapes in this
// process
std::mutex g_shape_mu;
std::map<std::pair<int, std::string>, int> g_shape_cache;

std::string gpu_model()
{
    int dev = 0;
    hipDeviceProp_t pr{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess) return "?";
    return std::string(pr.gcnArchName);
}

int cached_shape(int N)
{
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_shape_cache.find({N, gpu_model()});
    return it == g_shape_cache.end() ? 0 : it->second;
}

void set_member_variant(vk_ens* e, int N, int v)
{
    e->var[N] = v;
    for (int r : e->size_reals[N]) {
        vk_gpe* h = e->h[r];
        h->ens_var = v;
        int TC = 0;
        dispatch_var(N, v, [&](auto c) { TC = decltype(c)::TC; return 0; });
        h->geo_v = geo_of(N, TC);
    }
}

// (re)build the member table, the block maps of all launch modes and drop
// captured graphs (after a shape change)
int ens_rebuild(vk_ens* e)
{
    for (auto& g : e->gexec)
        if (g) { hipGraphExecDestroy(g); g = nullptr; }
    for (int k = 0; k < 3; ++k) {
        if (e->dmap[k]) { hipFree(e->dmap[k]); e->dmap[k] = nullptr; }
        if (e->doff[k]) { hipFree(e->doff[k]); e->doff[k] = nullptr; }
    }
    for (auto& kv : e->smaps) for (auto p : kv.second) if (p) hipFree(p);
    e->smaps.clear();
    e->snblk.clear();
    const int nreal = (int)e->h.size();
    std::vector<vk::EnsReal> R(nreal);
    for (int r = 0; r < nreal; ++r) R[r] = ens_entry(e->h[r]);
    if (!e->dR) HIPCHK(hipMalloc(&e->dR, sizeof(vk::EnsReal) * nreal));
    HIPCHK(hipMemcpy(e->dR, R.data(), sizeof(vk::EnsReal) * nreal, hipMemcpyHostToDevice));
    std::vector<int> all(nreal);
    for (int r = 0; r < nreal; ++r) all[r] = r;
    int rc = build_maps(e->h, all, e->dmap, e->doff, e->nblk);
    if (rc) return rc;
    for (int N : e->sizes) {
        std::array<int*, 6> m{};
        int* dm[3];
        int* dof[3];
        std::array<int, 3> nb{};
        if ((rc = build_maps(e->h, e->size_reals[N], dm, dof, nb.data()))) return rc;
        for (int k = 0; k < 3; ++k) { m[2 * k] = dm[k]; m[2 * k + 1] = dof[k]; }
        e->smaps[N] = m;
        e->snblk[N] = nb;
    }
    e->synced = 0;
    return 0;
}

// Hpsi = IFFT_y(K0) on the retained panels of shape (N, var)
int col_inv_var(vk_gpe* h, int var)
{
    return dispatch_var(h->N, var, [&](auto c) {
        using C = decltype(c);
        const Geo& g = var ? h->geo_v : h->geo;
        const size_t lds = (size_t)C::TC * C::COLRS * sizeof(double);
        set_lds(vk::k_col_inv<C>, lds);
        hipLaunchKernelGGL(vk::k_col_inv<C>, dim3(g.Pp + g.Pn), dim3(C::COLT), lds, h->stream, (const cplx*)h->K0,
                           h->Hpsi, (const cplx*)h->tw, g, 1, 0);
        return 0;
    });
}

// Diagnostics of the members hs (all of size C::N, all with Hpsi in layout
// pl: 1 retained panels, 0 all panels) as batched launches on stream st:
// the single-field pipeline (diag_launch) with blockIdx.y selecting the
// member, so each member gets bit for bit its single-field result.  Member
// i's sums go to sums[i] (8 doubles) and its spectra to spec[i].
// tab: device memory for the operand tables, ens_diag_tab_bytes(hs.size());
// htab: pinned host memory of the same size, not rewritten before the
// stream has passed the copy.
constexpr int kDiagPairSteps = 16;
size_t ens_diag_tab_bytes(size_t n)
{
    return ((kDiagPairSteps * n * sizeof(vk::BPair) + 255) / 256) * 256 + n * sizeof(vk::DiagB);
}

template <class C>
int ens_diag_group(const std::vector<vk_gpe*>& hs, int pl, hipStream_t st, void* tab, void* htab,
                   const std::vector<double*>& sums, const std::vector<double*>& spec)
{
    const int n = (int)hs.size();
    const vk_gpe* h0 = hs[0];
    const Geo g = h0->geo;
    const cplx* tw = h0->tw;   // the same table in every member of this size
    const double sc = scale_of(h0);
    const size_t dd_off = ((kDiagPairSteps * n * sizeof(vk::BPair) + 255) / 256) * 256;
    vk::BPair* pairs = static_cast<vk::BPair*>(htab);
    vk::DiagB* db = reinterpret_cast<vk::DiagB*>(static_cast<char*>(htab) + dd_off);
    for (int i = 0; i < n; ++i) {
        vk_gpe* h = hs[i];
        cplx** X = h->X;
        cplx** Kd = h->Kd;
        const vk::BPair row[kDiagPairSteps] = {
            {h->Hpsi, X[0]}, {h->Hpsi, X[1]}, {h->K0, h->Hd}, {h->Hd, X[2]},   // psi, d_x psi, d_y psi
            {X[0], h->Hd}, {h->Hd, Kd[4]},                                     // rho-hat, dealiased
            {Kd[4], h->Hd}, {h->Hd, X[3]},                                     // P_G rho
            {X[1], h->Hd}, {h->Hd, Kd[0]}, {X[2], h->Hd}, {h->Hd, Kd[1]},       // w, q transforms
            {X[5], h->Hd}, {h->Hd, Kd[2]}, {X[6], h->Hd}, {h->Hd, Kd[3]}};
        for (int k = 0; k < kDiagPairSteps; ++k) pairs[(size_t)k * n + i] = row[k];
        vk::DiagB& d = db[i];
        d.psi = X[0]; d.D1 = X[1]; d.D2 = X[2]; d.PR = X[3];
        d.W1 = X[1]; d.W2 = X[2]; d.Q1 = X[5]; d.Q2 = X[6];
        d.part = h->dpart;
        d.sums = sums[i];
        d.spec = spec[i];
        d.KW1 = Kd[0]; d.KW2 = Kd[1]; d.KQ1 = Kd[2]; d.KQ2 = Kd[3]; d.KR = Kd[4]; d.KP = h->K0;
        d.da = vk::DiagArgs{h->p.alpha, h->p.omega, h->p.g, (double)1e-30f, h->N, 0};
        const double en = 1.0 / ((double)h->N * (double)h->N);
        d.sa = vk::SpecArgs{h->p.alpha, h->p.g, en * en, h->N, h->nshell};
    }
    char* base = static_cast<char*>(tab);
    vk::BPair* dp = reinterpret_cast<vk::BPair*>(base);
    vk::DiagB* dd = reinterpret_cast<vk::DiagB*>(base + dd_off);
    HIPCHK(hipMemcpyAsync(tab, htab, ens_diag_tab_bytes(n), hipMemcpyHostToDevice, st));
    const size_t rl = (size_t)C::TR * C::ROWRS * sizeof(double), cl = (size_t)C::TC * C::COLRS * sizeof(double);
    set_lds(vk::k_row_inv_b<C>, rl);
    set_lds(vk::k_row_fwd_b<C>, rl);
    set_lds(vk::k_col_inv_b<C>, cl);
    set_lds(vk::k_col_fwd_b<C>, cl);
    const dim3 rg(C::N / C::TR, n), rb(C::ROWT), cb(C::COLT);
    const int ncl = pl == 1 ? g.Pp + g.Pn : C::NP;
    auto P = [&](int k) { return (const vk::BPair*)(dp + (size_t)k * n); };
    hipLaunchKernelGGL(vk::k_row_inv_b<C>, rg, rb, rl, st, P(0), tw, g, pl, 0, sc);
    hipLaunchKernelGGL(vk::k_row_inv_b<C>, rg, rb, rl, st, P(1), tw, g, pl, 1, sc);
    hipLaunchKernelGGL(vk::k_col_inv_b<C>, dim3(ncl, n), cb, cl, st, P(2), tw, g, pl, 2);
    hipLaunchKernelGGL(vk::k_row_inv_b<C>, rg, rb, rl, st, P(3), tw, g, pl, 0, sc);
    hipLaunchKernelGGL(vk::k_row_fwd_b<C>, rg, rb, rl, st, P(4), tw, g, 2, 1);
    hipLaunchKernelGGL(vk::k_col_fwd_b<C>, dim3(g.Pp, n), cb, cl, st, P(5), tw, g, 2, 1);
    hipLaunchKernelGGL(vk::k_col_inv_b<C>, dim3(g.Pp, n), cb, cl, st, P(6), tw, g, 2, 0);
    hipLaunchKernelGGL(vk::k_row_inv_b<C>, rg, rb, rl, st, P(7), tw, g, 2, 0, sc);
    const int nb = 1024;   // as the single-field pass: the same partial sums
    hipLaunchKernelGGL(vk::k_diag_x_b, dim3(nb, n), dim3(256), 0, st, (const vk::DiagB*)dd);
    hipLaunchKernelGGL(vk::k_reduce_fixed_b, dim3(n), dim3(256), 0, st, (const vk::DiagB*)dd, nb);
    for (int k = 8; k < kDiagPairSteps; k += 2) {
        hipLaunchKernelGGL(vk::k_row_fwd_b<C>, rg, rb, rl, st, P(k), tw, g, 2, 0);
        hipLaunchKernelGGL(vk::k_col_fwd_b<C>, dim3(g.Pp, n), cb, cl, st, P(k + 1), tw, g, 2, 0);
    }
    hipLaunchKernelGGL(vk::k_spectra_b, dim3(h0->nshell + 1, n), dim3(256), 0, st, (const int*)h0->modes,
                       (const int*)h0->offs, (const vk::DiagB*)dd);
    return 0;
}

}  // namespace

extern "C" {

int vk_ens_create(vk_ens** out, int nreal, const vk_gpe_params* params)
{
    if (!out || nreal <= 0 || !params) return fail(VK_ERR_ARG, "ens_create: bad arguments");
    *out = nullptr;
    for (int r = 0; r < nreal; ++r) {
        const int N = params[r].N;
        if (N != 128 && N != 256 && N != 512 && N != 1024)
            return fail(VK_ERR_UNSUPPORTED, "ensemble members must have N in {128, 256, 512, 1024}");
        if (params[r].np != 0) return fail(VK_ERR_UNSUPPORTED, "ensemble members carry no particles");
    }
    vk_ens* e = new vk_ens();
    auto cleanup = [&](int rc) { vk_ens_destroy(e); return rc; };
    for (int r = 0; r < nreal; ++r) {
        vk_gpe* h = nullptr;
        int rc = vk_gpe_create(&h, &params[r]);
        if (rc) return cleanup(rc);
        e->h.push_back(h);
        e->size_reals[h->N].push_back(r);
    }
    for (auto& kv : e->size_reals) e->sizes.push_back(kv.first);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(VK_ERR_HIP, "ens stream"));
    // LDS limits for every shape a member may take
    for (int N : e->sizes)
        for (int v = 0; v <= 1; ++v)
            if (vk::has_variant(N, v)) dispatch_var(N, v, [&](auto c) { ens_lds<decltype(c)>(e->lds); return 0; });
    set_lds(vk::k_ens_p1, e->lds[0]);
    set_lds(vk::k_ens_p2, e->lds[1]);
    set_lds(vk::k_ens_p3, e->lds[2]);
    set_lds(vk::k_ens_p4, e->lds[3]);
    for (auto k : {vk::k_ens_p4s<1>, vk::k_ens_p4s<2>, vk::k_ens_p4s<3>, vk::k_ens_p4s<4>}) set_lds(k, e->lds[3]);
    for (int N : e->sizes) dispatch(N, [&](auto n) { preconfig_passes<CfgN<decltype(n)::value>>(); return 0; });
    // the persistent step: as many workgroups as the device holds at once
    {
        e->plds = *std::max_element(e->lds, e->lds + 4);
        set_lds(vk::k_ens_step, e->plds);
        int dev = 0, per_cu = 0, ncu = 0, rate_khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, vk::k_ens_step, 256, e->plds) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "ens_create: device attributes"));
        e->pgrid = (per_cu * ncu) / 8 * 8;
        e->plimit = (long long)std::max(rate_khz, 1000) * 1000;   // one second
        if (hipMalloc(&e->dbar, sizeof(unsigned) * vk::kEnsStepBarriers) != hipSuccess ||
            hipMalloc(&e->derr, sizeof(int)) != hipSuccess || hipMemset(e->derr, 0, sizeof(int)) != hipSuccess)
            return cleanup(fail(VK_ERR_HIP, "ens_create: barrier memory"));
    }
    // shapes an earlier autotuning picked for this size on this GPU model
    for (int N : e->sizes) {
        const int v = cached_shape(N);
        if (v > 0) set_member_variant(e, N, v);
    }
    int rc = ens_rebuild(e);
    if (rc) return cleanup(rc);
    *out = e;
    return 0;
}

void vk_ens_destroy(vk_ens* e)
{
    if (!e) return;
    if (e->stream) hipStreamSynchronize(e->stream);
    for (auto g : e->gexec) if (g) hipGraphExecDestroy(g);
    for (int k = 0; k < 3; ++k) { if (e->dmap[k]) hipFree(e->dmap[k]); if (e->doff[k]) hipFree(e->doff[k]); }
    for (auto& kv : e->smaps) for (auto p : kv.second) if (p) hipFree(p);
    if (e->dR) hipFree(e->dR);
    if (e->dbar) hipFree(e->dbar);
    if (e->derr) hipFree(e->derr);
    if (e->hstage) hipHostFree(e->hstage);
    if (e->dstage) hipFree(e->dstage);
    if (e->dtab) hipFree(e->dtab);
    if (e->htab) hipHostFree(e->htab);
    for (auto h : e->h) vk_gpe_destroy(h);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

}  // extern "C"

namespace {

// After the stream has drained: fail if a persistent step's barrier wait ran
// out of time (a workgroup was not resident; that step's results are
// invalid).  The flag is cleared for later steps.
int ens_check_barrier(vk_ens* e)
{
    int err = 0;
    HIPCHK(hipMemcpy(&err, e->derr, sizeof(int), hipMemcpyDeviceToHost));
    if (!err) return 0;
    HIPCHK(hipMemset(e->derr, 0, sizeof(int)));
    return fail(VK_ERR_STATE, "ensemble mode 3: a grid barrier timed out (workgroups not co-resident); state invalid");
}

// one RK4 step of every realization on the ensemble stream
int ens_step_launch(vk_ens* e)
{
    if (e->mode == 2) {
        // per realization: the single-field kernels, one launch per pass each
        for (auto h : e->h) {
            hipStream_t keep = h->stream;
            h->stream = e->stream;
            int rc = rk4_step(h);
            h->stream = keep;
            if (rc) return rc;
        }
        return 0;
    }
    if (e->mode == 3) {
        if (e->pgrid < 8) return fail(VK_ERR_UNSUPPORTED, "ensemble: the device holds too few workgroups for mode 3");
        vk::EnsStepArgs A{};
        // one group per member size, blocks numbered over the sizes ascending
        if ((int)e->sizes.size() > vk::kEnsMaxGroups) return fail(VK_ERR_UNSUPPORTED, "ensemble: too many sizes");
        int base[3] = {};
        for (int N : e->sizes) {
            vk::EnsGroup& G = A.grp[A.ngrp++];
            G.key = N * 2 + (e->var.count(N) ? e->var[N] : 0);
            const auto& m = e->smaps[N];
            const auto& nb = e->snblk[N];
            for (int k = 0; k < 3; ++k) {
                G.map[k] = m[2 * k];
                G.off[k] = m[2 * k + 1];
                G.nblk[k] = nb[k];
                G.base[k] = base[k];
                base[k] += nb[k];
            }
        }
        A.bar = e->dbar;
        A.err = e->derr;
        A.limit = e->plimit;
        // no more workgroups than the largest pass has blocks: a small
        // ensemble's barriers then wait for few arrivals
        const int most = std::max(base[0], std::max(base[1], base[2]));
        const int grid = std::min(e->pgrid, (most + 7) / 8 * 8);
        hipLaunchKernelGGL(vk::k_zero_u32, dim3(1), dim3(64), 0, e->stream, e->dbar, vk::kEnsStepBarriers);
        hipLaunchKernelGGL(vk::k_ens_step, dim3(grid), dim3(256), e->plds, e->stream, (const vk::EnsReal*)e->dR, A);
        return 0;
    }
    struct Grp { int* m[3]; int* o[3]; int nb[3]; };
    std::vector<Grp> gs;
    if (e->mode == 0) {
        gs.push_back({{e->dmap[0], e->dmap[1], e->dmap[2]}, {e->doff[0], e->doff[1], e->doff[2]},
                      {e->nblk[0], e->nblk[1], e->nblk[2]}});
    } else {
        for (int N : e->sizes) {
            auto& m = e->smaps[N];
            auto& nb = e->snblk[N];
            gs.push_back({{m[0], m[2], m[4]}, {m[1], m[3], m[5]}, {nb[0], nb[1], nb[2]}});
        }
    }
    if (gs.size() == 1) {
        ens_passes(e->stream, e->dR, gs[0].m, gs[0].o, gs[0].nb, e->lds, nullptr);
        return 0;
    }
    // one launch per size, sizes interleaved pass by pass
    for (int stage = 1; stage <= 4; ++stage) {
        for (auto& g : gs)
            hipLaunchKernelGGL(vk::k_ens_p1, dim3(g.nb[0]), dim3(256), e->lds[0], e->stream,
                               (const vk::EnsReal*)e->dR, (const int*)g.m[0], (const int*)g.o[0]);
        for (auto& g : gs)
            hipLaunchKernelGGL(vk::k_ens_p2, dim3(g.nb[1]), dim3(256), e->lds[1], e->stream,
                               (const vk::EnsReal*)e->dR, (const int*)g.m[1], (const int*)g.o[1]);
        for (auto& g : gs)
            hipLaunchKernelGGL(vk::k_ens_p3, dim3(g.nb[0]), dim3(256), e->lds[2], e->stream,
                               (const vk::EnsReal*)e->dR, (const int*)g.m[0], (const int*)g.o[0]);
        for (auto& g : gs)
            ens_p4_launch(e->stream, g.nb[2], e->lds[3], (const vk::EnsReal*)e->dR, (const int*)g.m[2],
                          (const int*)g.o[2], stage);
    }
    return 0;
}

// Every member's Hpsi = IFFT_y(K0) in its shape's layout (the one-launch
// modes), or in the single-field layout (one launch per member).  A member
// of shape variant 1 then marks its Hpsi stale for the single-field entry
// points (hstate 0), which rebuild it in their layout on access.
int ens_prepare(vk_ens* e)
{
    if (e->synced) return 0;
    for (auto h : e->h) {
        int rc = ensure_rk4(h);
        if (rc) return rc;
        if (e->mode != 2 && h->ens_var) {
            if ((rc = col_inv_var(h, h->ens_var))) return rc;
            h->hstate = 0;
        }
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    e->synced = 1;
    return 0;
}

int ens_steps_async(vk_ens* e, int nsteps)
{
    int rc = ens_prepare(e);
    if (rc) return rc;
    if (e->use_graph) {
        hipGraphExec_t& ge = e->gexec[e->mode];
        if (!ge) {
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            rc = ens_step_launch(e);
            hipError_t ce = hipStreamEndCapture(e->stream, &g);
            if (rc) return rc;
            HIPCHK(ce);
            HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
        }
        for (int i = 0; i < nsteps; ++i) HIPCHK(hipGraphLaunch(ge, e->stream));
    } else {
        for (int i = 0; i < nsteps; ++i)
            if ((rc = ens_step_launch(e))) return rc;
    }
    return 0;
}

}  // namespace

extern "C" {

int vk_ens_size(vk_ens* e) { return e ? (int)e->h.size() : 0; }
void* vk_ens_stream(vk_ens* e) { return e ? (void*)e->stream : nullptr; }

static int ens_member(vk_ens* e, int r, vk_gpe** h)
{
    if (!e || r < 0 || r >= (int)e->h.size()) return fail(VK_ERR_ARG, "ensemble: bad realization index");
    HIPCHK(hipStreamSynchronize(e->stream));
    *h = e->h[r];
    return 0;
}

int vk_ens_set_psi(vk_ens* e, int r, const double* psi, int flags)
{
    vk_gpe* h;
    int rc = ens_member(e, r, &h);
    if (rc) return rc;
    e->synced = 0;
    return vk_gpe_set_psi(h, psi, flags);
}

int vk_ens_set_ic_vortices(vk_ens* e, int r, int nv, const double* xs, const double* ys, const int* qs, int nimg,
                           int flags)
{
    vk_gpe* h;
    int rc = ens_member(e, r, &h);
    if (rc) return rc;
    e->synced = 0;
    return vk_gpe_set_ic_vortices(h, nv, xs, ys, qs, nimg, flags);
}

int vk_ens_get_psi(vk_ens* e, int r, double* psi, double* kpsi)
{
    vk_gpe* h;
    int rc = ens_member(e, r, &h);
    if (rc) return rc;
    if (h->ens_var && psi) e->synced = 0;   // psi rebuilds Hpsi in the single-field layout
    return vk_gpe_get_psi(h, psi, kpsi);
}

int vk_ens_diagnostics(vk_ens* e, int r, double* scal, double* spec, double* pscal)
{
    vk_gpe* h;
    int rc = ens_member(e, r, &h);
    if (rc) return rc;
    if (h->ens_var) e->synced = 0;
    return vk_gpe_diagnostics(h, scal, spec, pscal);
}

int vk_ens_diagnostics_all(vk_ens* e, double* scal, double* spec)
{
    if (!e) return fail(VK_ERR_ARG, "ens_diagnostics_all: null");
    HIPCHK(hipStreamSynchronize(e->stream));
    const size_t nreal = e->h.size();
    std::vector<size_t> offs(nreal + 1, 0);
    for (size_t r = 0; r < nreal; ++r) offs[r + 1] = offs[r] + 8 + 6 * (size_t)(e->h[r]->nshell + 1);
    const size_t need = offs[nreal];
    if (need > e->nstage) {
        if (e->hstage) hipHostFree(e->hstage);
        if (e->dstage) hipFree(e->dstage);
        e->hstage = nullptr;
        e->dstage = nullptr;
        e->nstage = 0;
        HIPCHK(hipHostMalloc((void**)&e->hstage, need * sizeof(double)));
        HIPCHK(hipMalloc(&e->dstage, need * sizeof(double)));
        e->nstage = need;
    }
    for (auto h : e->h) {
        int rc = diag_setup(h);
        if (rc) return rc;
    }
    // groups of one size and one Hpsi state, batched launches on the
    // ensemble stream, each group's operand tables in its own part of dtab
    std::map<std::pair<int, int>, std::vector<int>> groups;
    for (size_t r = 0; r < nreal; ++r) groups[{e->h[r]->N, e->h[r]->hstate}].push_back((int)r);
    // per group: the ensure table (one pair per member), then the group's
    // tables; pinned staging in the same layout, so no copy waits
    auto ens_bytes = [](size_t n) { return (n * sizeof(vk::BPair) + 255) / 256 * 256; };
    size_t tabb = 0;
    for (auto& kv : groups)
        tabb += ens_bytes(kv.second.size()) + (ens_diag_tab_bytes(kv.second.size()) + 255) / 256 * 256;
    if (tabb > e->ntab) {
        if (e->dtab) hipFree(e->dtab);
        if (e->htab) hipHostFree(e->htab);
        e->dtab = nullptr;
        e->htab = nullptr;
        e->ntab = 0;
        HIPCHK(hipMalloc(&e->dtab, tabb));
        HIPCHK(hipHostMalloc(&e->htab, tabb));
        e->ntab = tabb;
    }
    size_t toff = 0;
    for (auto& kv : groups) {
        std::vector<vk_gpe*> hs;
        std::vector<double*> su, sp;
        for (int r : kv.second) {
            hs.push_back(e->h[r]);
            su.push_back(e->dstage + offs[r]);
            sp.push_back(e->dstage + offs[r] + 8);
        }
        char* etab = static_cast<char*>(e->dtab) + toff;
        char* ehtab = static_cast<char*>(e->htab) + toff;
        toff += ens_bytes(hs.size());
        char* tab = static_cast<char*>(e->dtab) + toff;
        char* htab = static_cast<char*>(e->htab) + toff;
        toff += (ens_diag_tab_bytes(hs.size()) + 255) / 256 * 256;
        int rc = dispatch(kv.first.first, [&](auto nn) {
            using C = CfgN<decltype(nn)::value>;
            if (kv.first.second == 0) {
                // Hpsi stale (a member in its second shape): IFFT_y(K0) on
                // all panels in the single-field layout, as ensure_any does
                vk::BPair* pr = reinterpret_cast<vk::BPair*>(ehtab);
                for (size_t i = 0; i < hs.size(); ++i) pr[i] = vk::BPair{hs[i]->K0, hs[i]->Hpsi};
                HIPCHK(hipMemcpyAsync(etab, ehtab, hs.size() * sizeof(vk::BPair), hipMemcpyHostToDevice, e->stream));
                const size_t cl = (size_t)C::TC * C::COLRS * sizeof(double);
                set_lds(vk::k_col_inv_b<C>, cl);
                hipLaunchKernelGGL(vk::k_col_inv_b<C>, dim3(C::NP, (int)hs.size()), dim3(C::COLT), cl, e->stream,
                                   (const vk::BPair*)etab, (const cplx*)hs[0]->tw, hs[0]->geo, 0, 0);
                for (auto h : hs) h->hstate = 1;
            }
            return ens_diag_group<C>(hs, hs[0]->hstate == 2 ? 1 : 0, e->stream, tab, htab, su, sp);
        });
        if (rc) return rc;
        for (auto h : hs)
            if (h->ens_var) e->synced = 0;   // Hpsi now in the single-field layout
    }
    HIPCHK(hipMemcpyAsync(e->hstage, e->dstage, need * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipGetLastError());
    const double dg[4] = {};
    size_t so = 0;
    for (size_t r = 0; r < nreal; ++r) {
        const vk_gpe* h = e->h[r];
        const size_t ns = 6 * (size_t)(h->nshell + 1);
        diag_final(h, e->hstage + offs[r], e->hstage + offs[r] + 8, dg, scal ? scal + 8 * r : nullptr,
                   spec ? spec + so : nullptr, nullptr);
        so += ns;
    }
    return 0;
}

void vk_ens_forget_shapes(void)
{
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shape_cache.clear();
}

int vk_ens_set_shape(vk_ens* e, int N, int var)
{
    if (!e || !e->size_reals.count(N) || !vk::has_variant(N, var))
        return fail(VK_ERR_ARG, "ens_set_shape: no members of that size, or no such shape for it");
    HIPCHK(hipStreamSynchronize(e->stream));
    set_member_variant(e, N, var);
    return ens_rebuild(e);
}

int vk_ens_get_shape(vk_ens* e, int N, int* var)
{
    if (!e || !var || !e->size_reals.count(N)) return fail(VK_ERR_ARG, "ens_get_shape: bad arguments");
    *var = e->var.count(N) ? e->var[N] : 0;
    return 0;
}

int vk_ens_rk4_steps_async(vk_ens* e, int nsteps)
{
    if (!e || nsteps < 0) return fail(VK_ERR_ARG, "ens_rk4_steps: bad arguments");
    return ens_steps_async(e, nsteps);
}

int vk_ens_rk4_steps(vk_ens* e, int nsteps)
{
    int rc = vk_ens_rk4_steps_async(e, nsteps);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipGetLastError());
    return ens_check_barrier(e);
}

int vk_ens_set_mode(vk_ens* e, int mode, int use_graph)
{
    if (!e || mode < 0 || mode > 3) return fail(VK_ERR_ARG, "ens_set_mode: bad arguments");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (mode != e->mode) e->synced = 0;   // the per-member mode uses the single-field layout
    e->mode = mode;
    e->use_graph = use_graph ? 1 : 0;
    return 0;
}

int vk_ens_get_mode(vk_ens* e, int* mode, int* use_graph)
{
    if (!e) return fail(VK_ERR_ARG, "ens_get_mode: null");
    if (mode) *mode = e->mode;
    if (use_graph) *use_graph = e->use_graph;
    return 0;
}

// Autotuner: times every (launch mode, graph) schedule over nsteps steps from
// the current state, restores the state, keeps the fastest.  ms_out[8]:
// ms per step for modes 0..3 x {graph, direct}.
int vk_ens_autotune(vk_ens* e, int nsteps, double* ms_out)
{
    if (!e || nsteps <= 0) return fail(VK_ERR_ARG, "ens_autotune: bad arguments");
    int rc = ens_prepare(e);
    if (rc) return rc;
    // save the state (kpsi and its y-inverse); K1/K2 are rewritten by stage 1
    std::vector<cplx*> save;
    for (auto h : e->h) {
        cplx* b = nullptr;
        HIPCHK(hipMalloc(&b, 2 * f2(h) * sizeof(cplx)));
        HIPCHK(hipMemcpyAsync(b, h->K0, f2(h) * sizeof(cplx), hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(b + f2(h), h->Hpsi, f2(h) * sizeof(cplx), hipMemcpyDeviceToDevice, e->stream));
        save.push_back(b);
    }
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    double best = 1e300;
    int bm = 0, bg = 1;
    for (int m = 0; m < 4; ++m)
        for (int gph = 1; gph >= 0; --gph) {
            if (m != e->mode) e->synced = 0;
            e->mode = m;
            e->use_graph = gph;
            // warm: capture, graph upload on the first replays
            if ((rc = ens_steps_async(e, 3))) return rc;
            HIPCHK(hipEventRecord(a, e->stream));
            if ((rc = ens_steps_async(e, nsteps))) return rc;
            HIPCHK(hipEventRecord(b, e->stream));
            HIPCHK(hipEventSynchronize(b));
            if (ens_check_barrier(e)) {   // mode 3 not co-resident here: not a candidate
                if (ms_out) ms_out[2 * m + (1 - gph)] = -1.0;
                continue;
            }
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, a, b));
            const double per = ms / nsteps;
            if (ms_out) ms_out[2 * m + (1 - gph)] = per;
            if (per < best) { best = per; bm = m; bg = gph; }
        }
    if (bm != e->mode) e->synced = 0;
    e->mode = bm;
    e->use_graph = bg;
    for (size_t i = 0; i < e->h.size(); ++i) {
        vk_gpe* h = e->h[i];
        HIPCHK(hipMemcpyAsync(h->K0, save[i], f2(h) * sizeof(cplx), hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(h->Hpsi, save[i] + f2(h), f2(h) * sizeof(cplx), hipMemcpyDeviceToDevice, e->stream));
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    for (auto p : save) hipFree(p);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return 0;
}

// Shape autotuner: for each member size with a second shape (8 points per
// thread instead of 16), times whole-ensemble steps in the current launch
// mode with either shape for that size (the other sizes as chosen so far,
// sizes ascending), restoring kpsi before every trial, and keeps the faster
// one.  The choice is remembered per (N, GPU model) for ensembles created
// later in the process.  vars_out[i], ms_out[2i], ms_out[2i+1]: the chosen
// shape and the ms/step of shapes 0 and 1 for the i-th distinct size
// (ascending; -1 where a size has one shape).
int vk_ens_autotune_shapes(vk_ens* e, int nsteps, int* vars_out, double* ms_out)
{
    if (!e || nsteps <= 0) return fail(VK_ERR_ARG, "ens_autotune_shapes: bad arguments");
    if (e->mode == 2) return fail(VK_ERR_STATE, "ens_autotune_shapes: the per-member mode has one shape");
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<cplx*> save;
    for (auto h : e->h) {
        cplx* p = nullptr;
        HIPCHK(hipMalloc(&p, f2(h) * sizeof(cplx)));
        HIPCHK(hipMemcpy(p, h->K0, f2(h) * sizeof(cplx), hipMemcpyDeviceToDevice));
        save.push_back(p);
    }
    auto restore = [&]() -> int {
        for (size_t i = 0; i < e->h.size(); ++i) {
            vk_gpe* h = e->h[i];
            HIPCHK(hipMemcpy(h->K0, save[i], f2(h) * sizeof(cplx), hipMemcpyDeviceToDevice));
            h->hstate = 0;
        }
        e->synced = 0;
        return 0;
    };
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int rc = 0;
    for (size_t i = 0; i < e->sizes.size() && !rc; ++i) {
        const int N = e->sizes[i];
        if (vars_out) vars_out[i] = 0;
        if (ms_out) ms_out[2 * i] = ms_out[2 * i + 1] = -1.0;
        if (!vk::has_variant(N, 1)) continue;
        double t[2] = {0.0, 0.0};
        for (int v = 0; v <= 1 && !rc; ++v) {
            set_member_variant(e, N, v);
            if ((rc = ens_rebuild(e)) || (rc = restore()) || (rc = ens_steps_async(e, 1))) break;
            if (hipEventRecord(a, e->stream) != hipSuccess) { rc = fail(VK_ERR_HIP, "event"); break; }
            if ((rc = ens_steps_async(e, nsteps))) break;
            if (hipEventRecord(b, e->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess) {
                rc = fail(VK_ERR_HIP, "event");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            t[v] = ms / nsteps;
        }
        if (rc) break;
        const int best = t[1] < t[0] ? 1 : 0;
        set_member_variant(e, N, best);
        if ((rc = ens_rebuild(e))) break;
        if (vars_out) vars_out[i] = best;
        if (ms_out) { ms_out[2 * i] = t[0]; ms_out[2 * i + 1] = t[1]; }
        std::lock_guard<std::mutex> lk(g_shape_mu);
        g_shape_cache[{N, gpu_model()}] = best;
    }
    if (!rc) rc = restore();
    HIPCHK(hipStreamSynchronize(e->stream));
    for (auto p : save) hipFree(p);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return rc;
}

}  // extern "C"

// single-field transform shape (defined here, after the ensemble helpers it uses)
extern "C" {

int vk_gpe_set_shape(vk_gpe* h, int var)
{
    if (!h || (var != 0 && var != 1)) return fail(VK_ERR_ARG, "set_shape: bad arguments");
    if (!h->e1R || !vk::has_variant(h->N, var) || h->p.np > 0)
        return fail(VK_ERR_UNSUPPORTED, "set_shape: shapes exist for N in {256, 512, 1024} without particles");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (var == h->svar) return 0;
    if (h->hstate == 3) h->hstate = 0;   // Hpsi in the old shape's layout
    h->svar = var;
    int TC = 0;
    dispatch_var(h->N, var, [&](auto c) { TC = decltype(c)::TC; return 0; });
    h->geo_s = var ? geo_of(h->N, TC) : h->geo;
    return make_single_table(h);
}

int vk_gpe_get_shape(vk_gpe* h, int* var)
{
    if (!h || !var) return fail(VK_ERR_ARG, "get_shape: bad arguments");
    *var = h->svar;
    return 0;
}

}  // extern "C"
