// Device kernels of the GPE step.  Every transform pass is a row pass
// (TR rows of N points along x per workgroup) or a column pass (TC columns
// of N points along y per workgroup); pointwise physics is fused into the
// passes so that each field round trip to HBM carries as much work as
// possible.  See vk_cfg.hpp for the layouts.
#pragma once
#include "vk_cfg.hpp"

// optional occupancy request for the pass kernels (waves per SIMD), and for
// the P2/P4 column passes alone
#ifdef VK_WAVES
#define VK_OCC __attribute__((amdgpu_waves_per_eu(VK_WAVES)))
#else
#define VK_OCC
#endif
#ifdef VK_COL_WAVES
#define VK_COCC __attribute__((amdgpu_waves_per_eu(VK_COL_WAVES)))
#else
#define VK_COCC VK_OCC
#endif
// P4 at 4 waves per SIMD (<= 128 VGPRs: two 512-thread workgroups per CU)
// is 9% faster than at the compiler's choice of 134-181 VGPRs
#ifndef VK_P4_WAVES
#define VK_P4_WAVES 4
#endif
#if VK_P4_WAVES > 0
#define VK_P4OCC __attribute__((amdgpu_waves_per_eu(VK_P4_WAVES)))
#else
#define VK_P4OCC VK_COCC
#endif
#ifdef VK_P1_WAVES
#define VK_P1OCC __attribute__((amdgpu_waves_per_eu(VK_P1_WAVES)))
#else
#define VK_P1OCC VK_OCC
#endif
// P3 with real-space particles keeps its factor f in registers too, not in
// 64 KB of LDS slots (P3 0.186 -> 0.184 ms; results bitwise the same)
#ifndef VK_P3_FREG_PART
#define VK_P3_FREG_PART 1
#endif
namespace vk { constexpr bool kP3FregPart = VK_P3_FREG_PART != 0; }
// P3 on the psi-row path without particles (PM = 0) at 4 waves per SIMD
// (<= 128 VGPRs): +0.8% pure-GPE steps/s with two rows per workgroup.  With
// particles the compiler's choice (231 VGPRs with one row per workgroup,
// two workgroups per CU) is kept: at 3 waves per SIMD (168 VGPRs, 188 B of
// spills) P3 takes 0.243 instead of 0.163 ms.
#ifndef VK_P3_WAVES
#define VK_P3_WAVES 4
#endif
#if VK_P3_WAVES > 0
#ifndef VK_P3_PART_WAVES
#define VK_P3_PART_WAVES 1
#endif
#define VK_P3OCC(PM) __attribute__((amdgpu_waves_per_eu((PM) == 0 ? VK_P3_WAVES : VK_P3_PART_WAVES)))
#else
#define VK_P3OCC(PM) VK_OCC
#endif

namespace vk {

// ----------------------------------------------------------------------------
// thread <-> data mapping helpers
// ----------------------------------------------------------------------------

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Row passes remap so that each XCD takes a contiguous run of row
// blocks: neighbouring row blocks share 128-byte lines of the panel layout,
// and those lines then stay in one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

#define VK_ROW_SETUP(C)                                                    \
    constexpr int N = C::N, S = C::S, T = C::T, TC = C::TC, TR = C::TR;    \
    constexpr int NP = C::NP;                                              \
    (void)NP;                                                              \
    extern __shared__ double lds[];                                        \
    const int u = tid_of<C>();                                             \
    const int q_ = u % (TR * TC);                                          \
    const int rr = q_ / TC, c = q_ % TC, pidx0 = u / (TR * TC);            \
    const int t = TC * pidx0 + c;                                          \
    const int y = xcd_remap(blk_, nblk_) * TR + rr;                       \
    double* lg = lds + rr * C::ROWRS;

#define VK_COL_SETUP(C, PEXPR)                                             \
    constexpr int N = C::N, S = C::S, T = C::T, TC = C::TC;                \
    constexpr int NP = C::NP;                                              \
    (void)NP;                                                              \
    extern __shared__ double lds[];                                        \
    const int u = tid_of<C>();                                             \
    const int c = u % TC, t = u / TC;                                      \
    const int p = (PEXPR);                                                 \
    const int kx = p * TC + c;                                             \
    double* lg = lds + c * C::COLRS;

// Column-pass setup for G columns per workgroup (G divides TC).  For
// G < TC the TC/G workgroups of a panel are siblings placed on one XCD:
// blocks come in groups of 8*TC/G covering 8 panels, and blocks b and b + 8
// (consecutive on one XCD under round-robin dispatch) share a panel.  Blocks
// past NPAN panels return here, before any barrier.
#define VK_COL_SETUP_G(C, G, B, PEXPR, NPAN)                                \
    constexpr int N = C::N, S = C::S, T = C::T, TC = C::TC;                \
    constexpr int NP = C::NP;                                              \
    (void)NP;                                                              \
    constexpr int GS_ = TC / (G);                                          \
    extern __shared__ double lds[];                                        \
    int pb_ = (B), hh_ = 0;                                                \
    if constexpr (GS_ > 1) {                                               \
        const int r_ = pb_ % (8 * GS_);                                    \
        hh_ = r_ / 8;                                                      \
        pb_ = pb_ / (8 * GS_) * 8 + r_ % 8;                                \
        if (pb_ >= (NPAN)) return;                                         \
    }                                                                      \
    const int p = (PEXPR);                                                 \
    const int u = tid_of<C>();                                             \
    const int cl_ = u % (G), t = u / (G);                                  \
    const int c = hh_ * (G) + cl_;                                         \
    const int kx = p * TC + c;                                             \
    double* lg = lds + cl_ * C::COLRS;

// panel holding kx = t + T*s in a row pass
#define VK_PANEL(s) (pidx0 + (T / TC) * (s))

__device__ __forceinline__ cplx czero() { return mk(0.0, 0.0); }

// ----------------------------------------------------------------------------
// Generic transform passes (set/get, diagnostics, conformance hook)
// ----------------------------------------------------------------------------

// x-space rows -> FFT along x -> H.  op 1: transform |X|^2 instead of X.
template <class C>
__device__ __forceinline__ void row_fwd_body(int blk_, int nblk_, const cplx* __restrict__ X, cplx* __restrict__ H,
                                             const cplx* __restrict__ tw, Geo g, int pstore, int op)
{
    VK_ROW_SETUP(C)
    cplx v[S];
    const cplx* xr = X + (size_t)y * N;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        cplx a = xr[t + T * s];
        v[s] = op == 1 ? mk(a.x * a.x + a.y * a.y, 0.0) : a;
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int pp = VK_PANEL(s);
        if (panel_stored(pp, NP, g, pstore)) H[hoff<C>(pp, y, c)] = v[s];
    }
}

template <class C>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_row_fwd(const cplx* __restrict__ X, cplx* __restrict__ H,
                                                     const cplx* __restrict__ tw, Geo g, int pstore, int op)
{
    row_fwd_body<C>(blockIdx.x, gridDim.x, X, H, tw, g, pstore, op);
}

// H column -> FFT along y -> K.  op 1: dealias.
// panels a column pass over pmode covers (0 all, 1 retained, 2 kx >= 0)
#define VK_NPAN(C, g, pmode) ((pmode) == 0 ? C::NP : ((pmode) == 2 ? (g).Pp : (g).Pp + (g).Pn))

template <class C, int G = C::TC>
__device__ __forceinline__ void col_fwd_body(int b, const cplx* __restrict__ H, cplx* __restrict__ K,
                                             const cplx* __restrict__ tw, Geo g, int pmode, int op)
{
    VK_COL_SETUP_G(C, G, b, panel_of_block(pb_, C::NP, g, pmode), VK_NPAN(C, g, pmode))
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = H[hoff<C>(p, t + T * s, c)];
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    const int kxs = ksg(kx, N);
    cplx* kr = K + (size_t)kx * N;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int ky = t + T * s;
        cplx a = v[s];
        if (op == 1 && kmasked(kxs, ksg(ky, N), N)) a = czero();
        kr[ky] = a;
    }
}

template <class C, int G = C::TC>
__global__ void __launch_bounds__(G * C::T) VK_OCC k_col_fwd(const cplx* __restrict__ H, cplx* __restrict__ K,
                                                     const cplx* __restrict__ tw, Geo g, int pmode, int op)
{
    col_fwd_body<C, G>(blockIdx.x, H, K, tw, g, pmode, op);
}

// K row -> op -> unnormalised inverse FFT along y -> H.
// op 0: as is; 1: dealias; 2: times i*ky.
template <class C, int G = C::TC>
__device__ __forceinline__ void col_inv_body(int b, const cplx* __restrict__ K, cplx* __restrict__ H,
                                             const cplx* __restrict__ tw, Geo g, int pmode, int op)
{
    VK_COL_SETUP_G(C, G, b, panel_of_block(pb_, C::NP, g, pmode), VK_NPAN(C, g, pmode))
    cplx v[S];
    const int kxs = ksg(kx, N);
    const cplx* kr = K + (size_t)kx * N;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int ky = t + T * s;
        const int kys = ksg(ky, N);
        cplx a = kr[ky];
        if (op == 1 && kmasked(kxs, kys, N)) a = czero();
        if (op == 2) a = mk(-(double)kys * a.y, (double)kys * a.x);
        v[s] = conjc(a);
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) H[hoff<C>(p, t + T * s, c)] = conjc(v[s]);
}

template <class C, int G = C::TC>
__global__ void __launch_bounds__(G * C::T) VK_OCC k_col_inv(const cplx* __restrict__ K, cplx* __restrict__ H,
                                                     const cplx* __restrict__ tw, Geo g, int pmode, int op)
{
    col_inv_body<C, G>(blockIdx.x, K, H, tw, g, pmode, op);
}

// H row -> op -> unnormalised inverse FFT along x -> times scale -> X.
// pload 2: kx >= 0 panels stored, negative kx from Hermitian symmetry.
// op 1: times i*kx before the transform.
template <class C>
__device__ __forceinline__ void row_inv_body(int blk_, int nblk_, const cplx* __restrict__ H, cplx* __restrict__ X,
                                             const cplx* __restrict__ tw, Geo g, int pload, int op, double scale)
{
    VK_ROW_SETUP(C)
    cplx v[S];
    const int Kp = TC * g.Pp;
    const brsrc rH = mkbuf(H, N);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int kxi = t + T * s;
        unsigned o;
        if (pload == 2) {
            const int km = kxi >= Kp ? N - kxi : kxi;
            o = boff(hoff<C>(km / TC, y, km % TC), kxi < Kp || kxi > N - Kp);
        } else {
            const int pp = VK_PANEL(s);
            o = boff(hoff<C>(pp, y, c), panel_stored(pp, NP, g, pload));
        }
        v[s] = bld(rH, o);
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int kxi = t + T * s;
        cplx a = v[s];
        if (pload == 2 && kxi > N - Kp) a = conjc(a);
        if (op == 1) {
            const double kk = (double)ksg(kxi, N);
            a = mk(-kk * a.y, kk * a.x);
        }
        v[s] = conjc(a);
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    cplx* xr = X + (size_t)y * N;
#pragma unroll
    for (int s = 0; s < S; ++s) xr[t + T * s] = rmul(scale, conjc(v[s]));
}

template <class C>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_row_inv(const cplx* __restrict__ H, cplx* __restrict__ X,
                                                     const cplx* __restrict__ tw, Geo g, int pload, int op,
                                                     double scale)
{
    row_inv_body<C>(blockIdx.x, gridDim.x, H, X, tw, g, pload, op, scale);
}

// zero the dealiased modes of a k-space field (all N^2 entries)
__global__ void k_dealias(cplx* __restrict__ K, int N)
{
    const size_t n2 = (size_t)N * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i / N), ky = (int)(i % N);
        if (kmasked(ksg(kx, N), ksg(ky, N), N)) K[i] = czero();
    }
}

// ----------------------------------------------------------------------------
// Fused RK4 / ARGLE passes
// ----------------------------------------------------------------------------

// P1 after its loads: v holds the conjugated Hpsi row points of this thread
template <class C, bool XPSI>
__device__ __forceinline__ void p1_core(cplx (&v)[C::S], double* lg, int t, int y, int pidx0, int c,
                                        cplx* __restrict__ Hrho, const cplx* __restrict__ tw, const Geo& g,
                                        double scale, cplx* __restrict__ Xpsi)
{
    constexpr int N = C::N, S = C::S, T = C::T, TC = C::TC;
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    if constexpr (XPSI) {
        cplx* xr = Xpsi + (size_t)y * N;
#pragma unroll
        for (int s = 0; s < S; ++s) xr[t + T * s] = rmul(scale, conjc(v[s]));
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double a = scale * v[s].x, b = scale * v[s].y;   // conj does not change |.|
        v[s] = mk(a * a + b * b, 0.0);
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int pp = VK_PANEL(s);
        if (pp < g.Pp) Hrho[hoff<C>(pp, y, c)] = v[s];
    }
}

// P1 (row): psi row from Hpsi (inverse along x, times 1/N^2), rho = |psi|^2,
// forward along x, keep kx >= 0 retained panels (rho is real).  XPSI: also
// store the psi row in x-space (Xpsi[y*N + x]) for P3, which then needs no
// transform of its own for psi.
template <class C, bool XPSI = false>
__device__ __forceinline__ void p1_body(int blk_, int nblk_, const cplx* __restrict__ Hpsi, cplx* __restrict__ Hrho,
                                                const cplx* __restrict__ tw, Geo g, int psi_pmode,
                                                double scale, cplx* __restrict__ Xpsi = nullptr)
{
    VK_ROW_SETUP(C)
    cplx v[S];
    const brsrc rH = mkbuf(Hpsi, N);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int pp = VK_PANEL(s);
        v[s] = bld(rH, boff(hoff<C>(pp, y, c), panel_stored(pp, NP, g, psi_pmode)));
    }
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = panel_stored(VK_PANEL(s), NP, g, psi_pmode) ? conjc(v[s]) : czero();
    p1_core<C, XPSI>(v, lg, t, y, pidx0, c, Hrho, tw, g, scale, Xpsi);
}

template <class C>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_p1(const cplx* __restrict__ Hpsi, cplx* __restrict__ Hrho,
                                                const cplx* __restrict__ tw, Geo g, int psi_pmode,
                                                double scale)
{
    p1_body<C>(blockIdx.x, gridDim.x, Hpsi, Hrho, tw, g, psi_pmode, scale);
}

template <class C>
__global__ void __launch_bounds__(C::ROWT) VK_P1OCC k_p1x(const cplx* __restrict__ Hpsi, cplx* __restrict__ Hrho,
                                                 const cplx* __restrict__ tw, Geo g, int psi_pmode,
                                                 double scale, cplx* __restrict__ Xpsi)
{
    p1_body<C, true>(blockIdx.x, gridDim.x, Hpsi, Hrho, tw, g, psi_pmode, scale, Xpsi);
}

// Retained-panel geometry of a size, at compile time (as the host computes
// it), and the LDS slots of the prefetching P1: the points s of a thread
// whose panel can be stored get a 16-byte slot per lane and wave; points
// whose stored lanes are disjoint share a slot (at 4096^2: 11 slots for the
// 12 point indices with stored panels).
struct P1Slots {
    int slot[16];
    int nslot;
};
template <class C> constexpr int geo_pp()
{
    constexpr int M = C::N / 3, pp = (M + 1 + C::TC - 1) / C::TC, pn = (M + C::TC - 1) / C::TC;
    return pp + pn > C::NP ? C::NP : pp;
}
template <class C> constexpr int geo_pn()
{
    constexpr int M = C::N / 3, pp = (M + 1 + C::TC - 1) / C::TC, pn = (M + C::TC - 1) / C::TC;
    return pp + pn > C::NP ? 0 : pn;
}
template <class C> constexpr P1Slots p1_slots()
{
    static_assert(C::S <= 16 && C::T / C::TC <= 128, "slot map sized for S <= 16, T/TC <= 128");
    P1Slots m{};
    unsigned long long occ[16][2] = {};
    m.nslot = 0;
    for (int s = 0; s < C::S; ++s) {
        unsigned long long b[2] = {0ull, 0ull};
        bool any = false;
        for (int q = 0; q < C::T / C::TC; ++q) {
            const int p = q + (C::T / C::TC) * s;
            if (p < geo_pp<C>() || p >= C::NP - geo_pn<C>()) {
                b[q / 64] |= 1ull << (q % 64);
                any = true;
            }
        }
        m.slot[s] = -1;
        if (!any) continue;
        for (int k = 0; k < m.nslot; ++k)
            if (!(occ[k][0] & b[0]) && !(occ[k][1] & b[1])) {
                m.slot[s] = k;
                break;
            }
        if (m.slot[s] < 0) m.slot[s] = m.nslot++;
        occ[m.slot[s]][0] |= b[0];
        occ[m.slot[s]][1] |= b[1];
    }
    return m;
}
template <class C> constexpr size_t p1p_lds()
{
    return (size_t)C::TR * C::ROWRS * sizeof(double) + (size_t)p1_slots<C>().nslot * C::ROWT * sizeof(cplx);
}

// P1 of an RK4 stage (retained panels) on the psi-row path, persistent: each
// workgroup walks row blocks of its XCD's contiguous range (XCD x = block %
// 8 owns [x R/8, (x+1) R/8), its workgroups stride through it) and loads the
// next block's Hpsi points into LDS with asynchronous global->LDS copies
// while it transforms the current one.  Every workgroup runs the same fixed
// number of iterations (the host checks that the blocks divide evenly).
template <class C>
__global__ void __launch_bounds__(C::ROWT) k_p1xp(const cplx* __restrict__ Hpsi, cplx* __restrict__ Hrho,
                                                  const cplx* __restrict__ tw, Geo g, double scale,
                                                  cplx* __restrict__ Xpsi)
{
    constexpr P1Slots SM = p1_slots<C>();
    constexpr int N = C::N, S = C::S, T = C::T, TC = C::TC, TR = C::TR, NP = C::NP;
    constexpr int NW = C::ROWT / 64;
    static_assert(C::ROWT % 64 == 0, "whole waves");
    extern __shared__ double lds[];
    const int u = threadIdx.x;
    const int q_ = u % (TR * TC);
    const int rr = q_ / TC, c = q_ % TC, pidx0 = u / (TR * TC);
    const int t = TC * pidx0 + c;
    double* lg = lds + rr * C::ROWRS;
    cplx* buf = reinterpret_cast<cplx*>(lds + TR * C::ROWRS);
    const int lane = u % 64, wv = u / 64;
    const int R8 = (N / TR) / 8, G8 = gridDim.x / 8;
    const int xcd = blockIdx.x % 8, j = blockIdx.x / 8;
    const int K = R8 / G8;
    auto prefetch = [&](int blk) {
        const int yy = blk * TR + rr;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (SM.slot[s] < 0) continue;
            const int pp = pidx0 + (T / TC) * s;
            if (pp < g.Pp || pp >= NP - g.Pn)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(Hpsi + hoff<C>(pp, yy, c)),
                    (__attribute__((address_space(3))) void*)(buf + (SM.slot[s] * NW + wv) * 64), 16, 0, 0);
        }
    };
    prefetch(xcd * R8 + j);
    for (int k = 0; k < K; ++k) {
        const int y = (xcd * R8 + j + k * G8) * TR + rr;
        // this wave's copies have landed (vmcnt), and every wave's (barrier)
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        cplx v[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int pp = pidx0 + (T / TC) * s;
            v[s] = (SM.slot[s] >= 0 && (pp < g.Pp || pp >= NP - g.Pn))
                       ? conjc(buf[(SM.slot[s] * NW + wv) * 64 + lane])
                       : czero();
        }
        // all reads are done before the next copies overwrite the slots
        __syncthreads();
        if (k + 1 < K) prefetch(xcd * R8 + j + (k + 1) * G8);
        p1_core<C, true>(v, lg, t, y, pidx0, c, Hrho, tw, g, scale, Xpsi);
    }
}

// P2 (column, kx >= 0 retained panels): forward along y, dealias
// (P_G), inverse along y, in place.  flags & 1: also the particle
// potential column HV = IFFT_y(KV); flags & 2: store the dealiased
// rho-hat row into K-space Krho (for the particle force).
// FL >= 0: the flags as a compile-time constant (flags is then ignored).
template <class C, int G = C::TC, int FL = -1>
__device__ __forceinline__ void p2_body(int blk_, int nblk_, cplx* __restrict__ Hrho, cplx* __restrict__ HV,
                                                const cplx* __restrict__ KV, cplx* __restrict__ Krho,
                                                const cplx* __restrict__ tw, Geo g, int flags_)
{
    const int flags = FL >= 0 ? FL : flags_;
    VK_COL_SETUP_G(C, G, blk_, pb_, g.Pp)
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = Hrho[hoff<C>(p, t + T * s, c)];
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int ky = t + T * s;
        const bool m = kmasked(kx, ksg(ky, N), N);
        if (m) v[s] = czero();
        else if (flags & 2) Krho[(size_t)kx * N + ky] = v[s];
        v[s] = conjc(v[s]);
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) Hrho[hoff<C>(p, t + T * s, c)] = conjc(v[s]);
    if (flags & 1) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ky = t + T * s;
            v[s] = kmasked(kx, ksg(ky, N), N) ? czero() : conjc(KV[(size_t)kx * N + ky]);
        }
        fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
        for (int s = 0; s < S; ++s) HV[hoff<C>(p, t + T * s, c)] = conjc(v[s]);
    }
}

template <class C, int G = C::TC, int FL = -1>
__global__ void __launch_bounds__(G * C::T) VK_COCC k_p2(cplx* __restrict__ Hrho, cplx* __restrict__ HV,
                                                const cplx* __restrict__ KV, cplx* __restrict__ Krho,
                                                const cplx* __restrict__ tw, Geo g, int flags)
{
    p2_body<C, G, FL>(blockIdx.x, gridDim.x, Hrho, HV, KV, Krho, tw, g, flags);
}

// deterministic workgroup sum of NV values per thread; result valid in
// thread 0.  sh must hold NV * (blockDim/64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&x)[NV], double* sh)
{
    if (blockDim.x < 64) {
        // partial wave: no cross-lane shuffles over inactive lanes
        const int nt = blockDim.x;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i * nt + threadIdx.x] = x[i];
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                double s = 0.0;
                for (int k = 0; k < nt; ++k) s += sh[i * nt + k];
                x[i] = s;
            }
        }
        __syncthreads();
        return;
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x[i] += __shfl_xor(x[i], o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i * nw + w] = x[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0.0;
            for (int k = 0; k < nw; ++k) s += sh[i * nw + k];
            x[i] = s;
        }
    }
    __syncthreads();
}

// Real-space particle terms for P3 (used when the particle potential's
// spectrum is negligible beyond the dealias radius, so that the reference's
// spectral translation IFFT(exp(-i k.q) Vhat) equals V(x - q) to far below
// FP64 roundoff; see the host-side selection).
struct P3Part {
    const double* pos;   // (x, y) per particle
    double* ppart;       // per row block, wave and particle: sum rho_raw * dV/dx, dV/dy
    unsigned char* phit; // per row block and particle: 1 if the window reaches the block
    int np;
    double V0, d, W, dx; // potential height, width, window radius, grid spacing
};

// minimum-image displacement in [-pi, pi)
__device__ __forceinline__ double pwrap(double a)
{
    const double tp = 6.283185307179586476925;
    return a - tp * rint(a * (1.0 / tp));
}

// periodic coordinate of src/gpe_particle.f90:46-50 (c = pi/L = 1/2):
// X = (1/c)(s + s^3/6 + s^5/40), s = sin(c a); also dX/da
__device__ __forceinline__ void pcoord(double a, double& X, double& dX)
{
    double sn, cs;
    sincos(0.5 * a, &sn, &cs);
    const double s2 = sn * sn;
    X = 2.0 * (sn + (1.0 / 6.0) * (s2 * sn) + (1.0 / 40.0) * (s2 * s2 * sn));
    dX = (1.0 + 0.5 * s2 + 0.125 * s2 * s2) * cs;
}

// Advective ARGLE terms (src/gpe.f90:78-81,91-94,114-146): tg adds the
// Taylor-Green velocity u = (sin x cos y, -cos x sin y) as
// -i u.grad psi - |u|^2 psi / (4 alpha); (vx, vy) the constant counterflow
// -i v.grad psi.  c4 = 1/(4 alpha); sc normalises the gradient transforms.
struct AdvArgs {
    int tg;
    double vx, vy, c4, sc;
    // a caller-supplied velocity field in place of the Taylor-Green one
    // (x-space, x fastest), when ux is set
    const double* ux;
    const double* uy;
};

// P3 (row): rho_d (+ i V) row from the kx >= 0 halves (Hermitian), psi row
// from Hpsi, N = (gs*rho_d + V) * psi, forward along x -> HN.
// gs = +-g/N^2, vs = +-1/N^2 (RK4 +, ARGLE -); psi scaled by 1/N^2.
// PM (particle mode, compile time): 0 none; 1 spectral potential (HV);
// 2 real-space potential; 6 real-space potential + force partial sums.
// vsign: +1 RK4, -1 ARGLE.
// XPSI: psi comes from the x-space rows P1 stored (already scaled), so the
// factor f stays in registers and no transform of Hpsi runs here.
// ADV: advective ARGLE; grad psi from the rows of i kx Hpsi and of Hdy =
// IFFT_y(i ky kpsi), two more transforms.
template <class C, int PM, bool XPSI = false, bool ADV = false>
__device__ __forceinline__ void p3_body(int blk_, int nblk_, const cplx* __restrict__ Hrhod, const cplx* __restrict__ HV,
                                                const cplx* __restrict__ Hpsi, cplx* __restrict__ HN,
                                                const cplx* __restrict__ tw, Geo g, int psi_pmode,
                                                int store_pmode, double gs, double vs, double scale,
                                                double vsign, P3Part pa, const cplx* __restrict__ Xpsi = nullptr,
                                                const cplx* __restrict__ Hdy = nullptr, AdvArgs aa = AdvArgs{})
{
    constexpr int flags = PM;
    VK_ROW_SETUP(C)
    cplx v[S];
    const int Kp = TC * g.Pp;
    // f in registers on the psi-row path (with real-space particles too,
    // unless VK_P3_FREG_PART=0), else parked in LDS slots after the
    // transform regions; particle positions staged after those during the
    // first transform, read back as broadcasts
    constexpr bool FREG = XPSI && ((flags & 2) == 0 || kP3FregPart);
    double* qlds = lds + TR * C::ROWRS + (FREG ? 0 : S * C::ROWT);
    if constexpr ((flags & 2) != 0) {
        for (int i = threadIdx.x; i < 2 * pa.np; i += blockDim.x) qlds[i] = pa.pos[i];
    }
    // kx >= 0 columns are stored; kx < 0 ones are the conjugates of N - kx
    const brsrc rR = mkbuf(Hrhod, N);
    cplx bv[(flags & 1) ? S : 1];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int kxi = t + T * s;
        const bool neg = kxi >= Kp;
        const int km = neg ? N - kxi : kxi;
        const unsigned o = boff(hoff<C>(km / TC, y, km % TC), kxi < Kp || kxi > N - Kp);
        v[s] = bld(rR, o);
        if constexpr ((flags & 1) != 0) bv[(flags & 1) ? s : 0] = bld(mkbuf(HV, N), o);
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const bool neg = t + T * s > N - Kp;
        cplx a = v[s], b = czero();
        if constexpr ((flags & 1) != 0) b = bv[(flags & 1) ? s : 0];
        if (neg) {
            a = conjc(a);
            b = conjc(b);
        }
        v[s] = conjc(mk(a.x - b.y, a.y + b.x));   // a + i b, conj for the inverse
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    // v[s].x = rho_d (unnormalised); v[s].y = -V (unnormalised, flags & 1).
    // The real factor f = g rho_d + V goes to thread-private LDS slots (after
    // the transform regions) and waits there while psi is transformed.
    double* fl = lds + TR * C::ROWRS;
    double fr_[FREG ? S : 1];
#define VK_F(s) (FREG ? fr_[FREG ? (s) : 0] : fl[(s) * C::ROWT + threadIdx.x])
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double fv;
        if constexpr ((flags & 1) != 0) fv = gs * v[s].x + vs * (-v[s].y);
        else fv = gs * v[s].x;
        VK_F(s) = fv;
    }
    if constexpr ((flags & 2) != 0) {
        // real-space particle potential added into the slots; force partials
        double rho[S];
#pragma unroll
        for (int s = 0; s < S; ++s) rho[s] = v[s].x;
        const int blk = xcd_remap(blk_, nblk_);
        const int yb = blk * TR;
        const double inv2d2 = 1.0 / (2.0 * (pa.d * pa.d)), invd2 = 1.0 / (pa.d * pa.d);
        const double idx = 1.0 / pa.dx, Wg = pa.W * idx, invN = 1.0 / N;
        // The particle loop has no workgroup barrier: every wave tests the
        // windows itself (one lane per particle, a ballot gives the hit set,
        // the same in every wave), reduces its force partials with lane
        // shuffles and writes them per wave; the per-block hit flags tell
        // the reduction which entries were written.
        constexpr int WV = C::ROWT < 64 ? C::ROWT : 64;   // lanes of a wave in use
        constexpr int NW = (C::ROWT + 63) / 64;            // waves per workgroup
        const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
        for (int base = 0; base < pa.np; base += WV) {
            // window test in grid units (multiplications only): minimum-image
            // row distance of the workgroup's rows to the particle
            const int pq = base + lane;
            bool hit = false;
            if (lane < WV && pq < pa.np) {
                const double cyq = qlds[2 * pq + 1] * idx;
                for (int r2 = 0; r2 < TR; ++r2) {
                    double dr = (yb + r2) - cyq;
                    dr -= N * rint(dr * invN);
                    hit |= fabs(dr) <= Wg;
                }
                if constexpr ((flags & 4) != 0)
                    if (wv == 0) pa.phit[(size_t)blk * pa.np + pq] = hit ? 1 : 0;
            }
            // hit particles in increasing order (uniform over the wave)
            for (unsigned long long m = __ballot(hit); m != 0ull; m &= m - 1ull) {
                const int pp = base + __builtin_ctzll(m);
                const double qx = qlds[2 * pp], qy = qlds[2 * pp + 1];
                const double cx = qx * idx, cy = qy * idx;   // centre in grid units
                double fr[2] = {0.0, 0.0};
                double dr = y - cy;
                dr -= N * rint(dr * invN);
                if (fabs(dr) <= Wg) {
                    double Y, dY;
                    pcoord(pwrap(y * pa.dx - qy), Y, dY);
                    // points of this thread inside the window (cheap test for
                    // all S), then the transcendental work once per such point,
                    // in increasing s (a single point when the window is
                    // narrower than the thread stride T)
                    unsigned mask = 0;
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        double dc = (t + T * s) - cx;
                        dc -= N * rint(dc * invN);
                        if (fabs(dc) <= Wg) mask |= 1u << s;
                    }
                    while (mask) {
                        const int sw = __builtin_ctz(mask);
                        mask &= mask - 1;
                        double rs = 0.0;
#pragma unroll
                        for (int s = 0; s < S; ++s) rs = (s == sw) ? rho[s] : rs;
                        double X, dX;
                        pcoord(pwrap((t + T * sw) * pa.dx - qx), X, dX);
                        const double e = pa.V0 * exp(-(X * X + Y * Y) * inv2d2);
                        if constexpr (FREG) {
#pragma unroll
                            for (int s = 0; s < S; ++s)
                                if (s == sw) fr_[FREG ? s : 0] += vsign * e;
                        } else {
                            fl[sw * C::ROWT + threadIdx.x] += vsign * e;
                        }
                        fr[0] += rs * (-e * X * dX * invd2);
                        fr[1] += rs * (-e * Y * dY * invd2);
                    }
                }
                if constexpr ((flags & 4) != 0) {
#pragma unroll
                    for (int o = WV / 2; o > 0; o >>= 1) {
                        fr[0] += __shfl_xor(fr[0], o, WV);
                        fr[1] += __shfl_xor(fr[1], o, WV);
                    }
                    if (lane == 0) {
                        double* q = pa.ppart + (((size_t)blk * NW + wv) * pa.np + pp) * 2;
                        q[0] = fr[0];
                        q[1] = fr[1];
                    }
                }
            }
        }
    }
    // Taylor-Green u at (x, y) from the twiddle table: tw[k] = e^{-2 pi i k/N}
    const double cy_ = tw[y].x, sy_ = -tw[y].y;
    if constexpr (ADV) {
        if (aa.ux) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const size_t i = (size_t)y * N + t + T * s;
                const double ux = aa.ux[i], uy = aa.uy[i];
                VK_F(s) -= aa.c4 * (ux * ux + uy * uy);
            }
        } else if (aa.tg) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const cplx e = tw[t + T * s];
                const double ux = -e.y * cy_, uy = -e.x * sy_;
                VK_F(s) -= aa.c4 * (ux * ux + uy * uy);
            }
        }
    }
    if constexpr (XPSI) {
        const cplx* xr = Xpsi + (size_t)y * N;
#pragma unroll
        for (int s = 0; s < S; ++s) v[s] = rmul(VK_F(s), xr[t + T * s]);
    } else {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int pp = VK_PANEL(s);
            v[s] = panel_stored(pp, NP, g, psi_pmode) ? conjc(Hpsi[hoff<C>(pp, y, c)]) : czero();
        }
        fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const cplx ps = rmul(scale, conjc(v[s]));
            v[s] = rmul(VK_F(s), ps);
        }
    }
#undef VK_F
    if constexpr (ADV) {
        // -i (u + v).grad psi, one gradient component at a time
        cplx w[S];
#pragma unroll 1
        for (int comp = 0; comp < 2; ++comp) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int pp = VK_PANEL(s);
                cplx a = czero();
                if (panel_stored(pp, NP, g, psi_pmode)) {
                    if (comp == 0) {
                        a = Hpsi[hoff<C>(pp, y, c)];
                        const double kk = (double)ksg(t + T * s, N);
                        a = mk(-kk * a.y, kk * a.x);   // i kx
                    } else {
                        a = Hdy[hoff<C>(pp, y, c)];
                    }
                }
                w[s] = conjc(a);
            }
            fft_fwd<N, S, C::SPLIT>(w, lg, t, tw);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const cplx e = tw[t + T * s];
                double uf;
                if (aa.ux) uf = (comp == 0 ? aa.ux : aa.uy)[(size_t)y * N + t + T * s];
                else uf = aa.tg ? (comp == 0 ? -e.y * cy_ : -e.x * sy_) : 0.0;
                const double u = uf + (comp == 0 ? aa.vx : aa.vy);
                const cplx d = rmul(aa.sc, conjc(w[s]));   // d psi / dx_comp
                v[s] = cadd(v[s], mk(u * d.y, -u * d.x));  // -i u d
            }
        }
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    // branch-free buffer stores (an unstored panel gets the out-of-range
    // offset and is dropped): conditional stores made the compiler sink the
    // last butterfly into one branch per store (P3 0.191 -> 0.187 ms)
    const brsrc rN = mkbuf(HN, N);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int pp = VK_PANEL(s);
        bst(rN, boff(hoff<C>(pp, y, c), panel_stored(pp, NP, g, store_pmode)), v[s]);
    }
}

template <class C, int PM>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_p3(const cplx* __restrict__ Hrhod, const cplx* __restrict__ HV,
                                                const cplx* __restrict__ Hpsi, cplx* __restrict__ HN,
                                                const cplx* __restrict__ tw, Geo g, int psi_pmode,
                                                int store_pmode, double gs, double vs, double scale,
                                                double vsign, P3Part pa)
{
    p3_body<C, PM>(blockIdx.x, gridDim.x, Hrhod, HV, Hpsi, HN, tw, g, psi_pmode, store_pmode, gs, vs, scale, vsign, pa);
}

template <class C, int PM, bool XPSI>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_p3a(const cplx* __restrict__ Hrhod, const cplx* __restrict__ HV,
                                                 const cplx* __restrict__ Hpsi, const cplx* __restrict__ Xpsi,
                                                 const cplx* __restrict__ Hdy, cplx* __restrict__ HN,
                                                 const cplx* __restrict__ tw, Geo g, double gs, double vs,
                                                 double scale, double vsign, P3Part pa, AdvArgs aa)
{
    p3_body<C, PM, XPSI, true>(blockIdx.x, gridDim.x, Hrhod, HV, Hpsi, HN, tw, g, 0, 0, gs, vs, XPSI ? 1.0 : scale,
                               vsign, pa, Xpsi, Hdy, aa);
}

template <class C, int PM>
__global__ void __launch_bounds__(C::ROWT) VK_P3OCC(PM) k_p3x(const cplx* __restrict__ Hrhod, const cplx* __restrict__ HV,
                                                 const cplx* __restrict__ Xpsi, cplx* __restrict__ HN,
                                                 const cplx* __restrict__ tw, Geo g, int store_pmode, double gs,
                                                 double vs, double vsign, P3Part pa)
{
    p3_body<C, PM, true>(blockIdx.x, gridDim.x, Hrhod, HV, nullptr, HN, tw, g, 0, store_pmode, gs, vs, 1.0, vsign,
                         pa, Xpsi);
}

struct RKArgs {
    cplx pref;      // (1 - i disp)/i
    double alpha, omega;
    double wacc;    // dt/6 (stages 1,4) or dt/6*2 (stages 2,3)
    double cstep;   // dt/2, dt/2, dt, - for stages 1..3
    double r1;      // stage 2: str2 after stage 1 = r1 (K2 - K0), r1 = (dt/6)/(dt/2)
    double r3;      // stage 4: str2 after stage 3 = K1 + r3 (K2 - K0), r3 = (dt/3)/dt
    int stage;
};

// P4 (column, retained panels both signs): forward along y -> N-hat;
// RHS = pref*((alpha k^2 - omega) psi_s-hat + N-hat) on the retained disk;
// RK4 accumulation and stage update (src/gpe.f90:366-487);
// inverse along y of the new kpsi -> Hpsi.
// K0 = str1 (= kpsi at step start), K1 = str2, K2 = stage kpsi.  Stages 1
// and 3 do not store str2: the term they would add, (dt/6) rhs1 or
// (dt/3) rhs3, equals r (K2 - K0) with the K2 they write, and the next
// stage rebuilds it from arrays it reads anyway (two fewer field writes
// and one fewer field read per step; each costs one rounding of K2, ~1 ulp
// of kpsi).  The stage is a compile-time constant, and the K accesses of
// the dealiased modes go to the out-of-range offset (loads read zero,
// stores are dropped), so the loads of all points are issued back to back.
// G < TC: G of a panel's TC columns per workgroup (sibling workgroups of a
// panel on one XCD, as in P2).
template <class C, int STAGE, int G = C::TC>
__device__ __forceinline__ void p4_body(int blk_, int nblk_, const cplx* __restrict__ HN, cplx* __restrict__ Hpsi,
                                                cplx* __restrict__ K0, cplx* __restrict__ K1,
                                                cplx* __restrict__ K2, const cplx* __restrict__ tw,
                                                Geo g, RKArgs a)
{
    VK_COL_SETUP_G(C, G, blk_, panel_of_block(pb_, C::NP, g, 1), g.Pp + g.Pn)
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = HN[hoff<C>(p, t + T * s, c)];
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    const int kxs = ksg(kx, N);
    const size_t rowo = (size_t)kx * N;
    const brsrc r0 = mkbuf(K0, N), r1 = mkbuf(K1, N), r2 = mkbuf(K2, N);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int ky = t + T * s;
        const int kys = ksg(ky, N);
        const bool keep = !kmasked(kxs, kys, N);
        const unsigned o = boff(rowo + ky, keep);
        const double kr2 = (double)(kxs * kxs + kys * kys);
        const cplx s1 = bld(r0, o);
        const cplx ps = STAGE == 1 ? s1 : bld(r2, o);
        const cplx lin = rmul(a.alpha * kr2 - a.omega, ps);
        const cplx rhs = cmul(a.pref, cadd(lin, v[s]));
        const cplx inc = rmul(a.wacc, rhs);
        cplx acc = inc, kn;
        if constexpr (STAGE == 2) acc = cadd(rmul(a.r1, csub(ps, s1)), inc);
        else if constexpr (STAGE == 3) acc = cadd(bld(r1, o), inc);
        else if constexpr (STAGE == 4) acc = cadd(cadd(bld(r1, o), rmul(a.r3, csub(ps, s1))), inc);
        if constexpr (STAGE < 4) {
            if constexpr (STAGE == 2) bst(r1, o, acc);
            kn = cadd(s1, rmul(a.cstep, rhs));
            bst(r2, o, kn);
        } else {
            kn = cadd(s1, acc);
            bst(r0, o, kn);
        }
        v[s] = conjc(keep ? kn : czero());
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) Hpsi[hoff<C>(p, t + T * s, c)] = conjc(v[s]);
}

template <class C, int STAGE, int G = C::TC>
__global__ void __launch_bounds__(G * C::T) VK_P4OCC k_p4(const cplx* __restrict__ HN, cplx* __restrict__ Hpsi,
                                                cplx* __restrict__ K0, cplx* __restrict__ K1,
                                                cplx* __restrict__ K2, const cplx* __restrict__ tw,
                                                Geo g, RKArgs a)
{
    p4_body<C, STAGE, G>(blockIdx.x, gridDim.x, HN, Hpsi, K0, K1, K2, tw, g, a);
}

struct ArgleArgs {
    double alpha, omega, dt;
    int want_res;
    // SGLE (src/gpe.f90:200-217): with Kn set, sqrt(dt) Kn (dealiased) is
    // added to the numerator, omega is read from *omega_dev (the chemical
    // potential of this step) and, for ktrunc2 > 0, modes with
    // k^2 > ktrunc2 are zeroed
    const cplx* Kn;
    const double* omega_dev;
    double sdt, ktrunc2;
};

// ARGLE A4 (column, all panels): forward along y -> N-hat; implicit update
// kpsi = (kpsi + dt*(N-hat + omega*kpsi))/(1 + alpha*k^2*dt) on every mode
// (src/gpe.f90:195-197, no dealiasing of the result); residual partial
// sums; inverse along y of the new kpsi -> Hpsi.  G < TC: G of a panel's
// columns per workgroup (the host keeps NP a multiple of 8 then, so that
// every workgroup writes its residual partials).
template <class C, int G = C::TC>
__device__ __forceinline__ void a4_body(int blk_, int nblk_, const cplx* __restrict__ HN, cplx* __restrict__ Hpsi,
                                                cplx* __restrict__ K0, const cplx* __restrict__ tw,
                                                ArgleArgs a, double* __restrict__ part)
{
    VK_COL_SETUP_G(C, G, blk_, pb_, C::NP)
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = HN[hoff<C>(p, t + T * s, c)];
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
    const int kxs = ksg(kx, N);
    const size_t rowo = (size_t)kx * N;
    double r[2] = {0.0, 0.0};
    const double omega = a.omega_dev ? *a.omega_dev : a.omega;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int ky = t + T * s;
        const int kys = ksg(ky, N);
        const size_t o = rowo + ky;
        const cplx old = K0[o];
        const double kr2 = (double)(kxs * kxs + kys * kys);
        const cplx t1 = cadd(v[s], rmul(omega, old));
        const double t2 = 1.0 + a.alpha * kr2 * a.dt;
        cplx num = cadd(old, rmul(a.dt, t1));
        if (a.Kn) {
            const cplx nt = kmasked(kxs, kys, N) ? czero() : rmul(a.sdt, a.Kn[o]);
            num = cadd(num, nt);
        }
        cplx nw = mk(num.x / t2, num.y / t2);
        if (a.ktrunc2 > 0.0 && kr2 > a.ktrunc2) nw = czero();
        K0[o] = nw;
        if (a.want_res) {
            const double dx = nw.x - old.x, dy = nw.y - old.y;
            r[0] += dx * dx + dy * dy;
            r[1] += old.x * old.x + old.y * old.y;
        }
        v[s] = conjc(nw);
    }
    fft_fwd<N, S, C::SPLIT>(v, lg, t, tw);
#pragma unroll
    for (int s = 0; s < S; ++s) Hpsi[hoff<C>(p, t + T * s, c)] = conjc(v[s]);
    if (a.want_res) {
        block_sum<2>(r, lds);
        if (threadIdx.x == 0) {
            part[2 * blk_] = r[0];
            part[2 * blk_ + 1] = r[1];
        }
    }
}

template <class C, int G = C::TC>
__global__ void __launch_bounds__(G * C::T) VK_OCC k_a4(const cplx* __restrict__ HN, cplx* __restrict__ Hpsi,
                                                cplx* __restrict__ K0, const cplx* __restrict__ tw,
                                                ArgleArgs a, double* __restrict__ part)
{
    a4_body<C, G>(blockIdx.x, gridDim.x, HN, Hpsi, K0, tw, a, part);
}

// fixed-order reduction of nb partial records of nv doubles -> out[nv]
__device__ __forceinline__ void reduce_fixed_body(const double* __restrict__ part, int nb, int nv,
                                                  double* __restrict__ out)
{
    __shared__ double sh[1024];
    for (int i = 0; i < nv; ++i) {
        double x = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) x += part[(size_t)b * nv + i];
        double xx[1] = {x};
        block_sum<1>(xx, sh);
        if (threadIdx.x == 0) out[i] = xx[0];
        __syncthreads();
    }
}

__global__ void k_reduce_fixed(const double* __restrict__ part, int nb, int nv, double* __restrict__ out)
{
    reduce_fixed_body(part, nb, nv, out);
}

// ----------------------------------------------------------------------------
// Diagnostics (src/diagnostics.f90:54-426)
// ----------------------------------------------------------------------------

struct DiagArgs {
    double alpha, omega, g, smlnum;
    int N;
    int has_v;
};

// x-space pass: momentum, energies, norm partial sums (in the reference's
// :84-207 quantities before the final factors) and the fields
//   w = 2 alpha Im(psi* grad psi)/(|psi|+smlnum) -> W1, W2
//   q = Re(psi* grad psi)/(|psi|+smlnum)          -> Q1, Q2   (:263-317)
// sums: 0 momx 1 momy 2 egrad 3 echem 4 eintr 5 eppot 6 norm 7 unused
__device__ __forceinline__ void diag_x_body(int blk, int nblk, const cplx* __restrict__ psi,
                                            const cplx* __restrict__ D1, const cplx* __restrict__ D2,
                                            const cplx* __restrict__ PR, const cplx* __restrict__ V,
                                            cplx* __restrict__ W1, cplx* __restrict__ W2, cplx* __restrict__ Q1,
                                            cplx* __restrict__ Q2, DiagArgs a, double* __restrict__ part)
{
    __shared__ double sh[8 * 16];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const size_t n2 = (size_t)a.N * a.N;
    for (size_t i = blk * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)nblk * blockDim.x) {
        const cplx ps = psi[i], d1 = D1[i], d2 = D2[i], pr = PR[i];
        const cplx cp = conjc(ps);
        const cplx a1 = cmul(cp, d1), a2 = cmul(cp, d2);
        const double apsi = hypot(ps.x, ps.y);
        const double ap2 = apsi * apsi;
        acc[0] += a1.y;
        acc[1] += a2.y;
        const double h1 = hypot(d1.x, d1.y), h2 = hypot(d2.x, d2.y);
        acc[2] += a.alpha * (h1 * h1 + h2 * h2);
        acc[3] += a.omega * ap2;
        acc[4] += 0.5 * a.g * hypot(pr.x, pr.y) * ap2;
        if (a.has_v) acc[5] += V[i].x * pr.x;
        acc[6] += ap2;
        const double rl = apsi + a.smlnum;
        W1[i] = mk(2.0 * a.alpha * (a1.y / rl), 0.0);
        W2[i] = mk(2.0 * a.alpha * (a2.y / rl), 0.0);
        Q1[i] = mk(a1.x / rl, 0.0);
        Q2[i] = mk(a2.x / rl, 0.0);
    }
    block_sum<8>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) part[(size_t)blk * 8 + k] = acc[k];
}

__global__ void k_diag_x(const cplx* __restrict__ psi, const cplx* __restrict__ D1,
                         const cplx* __restrict__ D2, const cplx* __restrict__ PR,
                         const cplx* __restrict__ V, cplx* __restrict__ W1, cplx* __restrict__ W2,
                         cplx* __restrict__ Q1, cplx* __restrict__ Q2, DiagArgs a, double* __restrict__ part)
{
    diag_x_body(blockIdx.x, gridDim.x, psi, D1, D2, PR, V, W1, W2, Q1, Q2, a, part);
}

struct SpecArgs {
    double alpha, g, enorm;
    int N, nshell;
};

// Shell binning: one workgroup per shell m over the list of retained modes
// in that shell (idx = kx*N + ky), fixed-order reduction.
// spec[(i)*(nshell+1) + m], i = 0..5 = spec_gp(m, i+1) (:394-399).
__device__ __forceinline__ void spectra_body(int m, const int* __restrict__ modes, const int* __restrict__ offs,
                                             const cplx* __restrict__ KW1, const cplx* __restrict__ KW2,
                                             const cplx* __restrict__ KQ1, const cplx* __restrict__ KQ2,
                                             const cplx* __restrict__ KR, const cplx* __restrict__ KP, SpecArgs a,
                                             double* __restrict__ spec)
{
    __shared__ double sh[6 * 16];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int j = offs[m] + threadIdx.x; j < offs[m + 1]; j += blockDim.x) {
        const int idx = modes[j];
        const int kx = ksg(idx / a.N, a.N), ky = ksg(idx % a.N, a.N);
        // the real fields' transforms hold the kx >= 0 columns: a mode with
        // kx < 0 takes the values of its mirror -k, whose terms are the same
        // (every term is a squared modulus, invariant under k -> -k with
        // the conjugated amplitudes)
        const int kxm = kx < 0 ? -kx : kx, kym = kx < 0 ? -ky : ky;
        const int idh = kxm * a.N + (kym < 0 ? kym + a.N : kym);
        const double kxd = kxm, kyd = kym, kr2 = kxd * kxd + kyd * kyd;
        const cplx aa1 = KW1[idh], aa2 = KW2[idh];
        cplx bb = czero();
        if (!(kx == 0 && ky == 0)) {
            // bb = -i (kx/k^2) aa1 - i (ky/k^2) aa2
            const double f1 = kxd / kr2, f2 = kyd / kr2;
            const cplx b1 = mk(f1 * aa1.y, -f1 * aa1.x), b2 = mk(f2 * aa2.y, -f2 * aa2.x);
            bb = cadd(b1, b2);
        }
        const cplx cc1 = mk(-kxd * bb.y, kxd * bb.x), cc2 = mk(-kyd * bb.y, kyd * bb.x);
        const cplx dd1 = csub(aa1, cc1), dd2 = csub(aa2, cc2);
        const cplx e1 = KQ1[idh], e2 = KQ2[idh], ff = KR[idh], gg = KP[idx];
        auto ab2 = [](cplx z) { const double h = hypot(z.x, z.y); return h * h; };
        acc[0] += a.enorm * (ab2(aa1) + ab2(aa2)) * 0.5;
        acc[1] += a.enorm * (ab2(dd1) + ab2(dd2)) * 0.5;
        acc[2] += a.enorm * (ab2(cc1) + ab2(cc2)) * 0.5;
        acc[3] += a.enorm * (ab2(e1) + ab2(e2)) * 2.0 * (a.alpha * a.alpha);
        acc[4] += a.enorm * ab2(ff) * 0.5 * a.g * 2.0 * a.alpha;
        acc[5] += a.enorm * ab2(gg);
    }
    block_sum<6>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 6; ++i) spec[(size_t)i * (a.nshell + 1) + m] = acc[i];
}

__global__ void k_spectra(const int* __restrict__ modes, const int* __restrict__ offs,
                          const cplx* __restrict__ KW1, const cplx* __restrict__ KW2,
                          const cplx* __restrict__ KQ1, const cplx* __restrict__ KQ2,
                          const cplx* __restrict__ KR, const cplx* __restrict__ KP, SpecArgs a,
                          double* __restrict__ spec)
{
    spectra_body(blockIdx.x, modes, offs, KW1, KW2, KQ1, KQ2, KR, KP, a, spec);
}

// ----------------------------------------------------------------------------
// Batched forms for many fields of one size (the ensemble's diagnostics):
// blockIdx.y selects the field's pointers, blockIdx.x is the block within
// the field, as in the single-field launch
// ----------------------------------------------------------------------------

struct BPair {
    const cplx* src;
    cplx* dst;
};

template <class C>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_row_fwd_b(const BPair* __restrict__ P, const cplx* __restrict__ tw,
                                                       Geo g, int pstore, int op)
{
    const BPair q = P[blockIdx.y];
    row_fwd_body<C>(blockIdx.x, gridDim.x, q.src, q.dst, tw, g, pstore, op);
}

template <class C>
__global__ void __launch_bounds__(C::COLT) VK_OCC k_col_fwd_b(const BPair* __restrict__ P, const cplx* __restrict__ tw,
                                                       Geo g, int pmode, int op)
{
    const BPair q = P[blockIdx.y];
    col_fwd_body<C>(blockIdx.x, q.src, q.dst, tw, g, pmode, op);
}

template <class C>
__global__ void __launch_bounds__(C::COLT) VK_OCC k_col_inv_b(const BPair* __restrict__ P, const cplx* __restrict__ tw,
                                                       Geo g, int pmode, int op)
{
    const BPair q = P[blockIdx.y];
    col_inv_body<C>(blockIdx.x, q.src, q.dst, tw, g, pmode, op);
}

template <class C>
__global__ void __launch_bounds__(C::ROWT) VK_OCC k_row_inv_b(const BPair* __restrict__ P, const cplx* __restrict__ tw,
                                                       Geo g, int pload, int op, double scale)
{
    const BPair q = P[blockIdx.y];
    row_inv_body<C>(blockIdx.x, gridDim.x, q.src, q.dst, tw, g, pload, op, scale);
}

// one field's diagnostics operands
struct DiagB {
    const cplx *psi, *D1, *D2, *PR;
    cplx *W1, *W2, *Q1, *Q2;
    double* part;     // 8 partial sums per block
    double* sums;     // 8 sums
    double* spec;     // 6 (nshell + 1) spectra
    const cplx *KW1, *KW2, *KQ1, *KQ2, *KR, *KP;
    DiagArgs da;
    SpecArgs sa;
};

__global__ void k_diag_x_b(const DiagB* __restrict__ D)
{
    const DiagB& d = D[blockIdx.y];
    diag_x_body(blockIdx.x, gridDim.x, d.psi, d.D1, d.D2, d.PR, nullptr, d.W1, d.W2, d.Q1, d.Q2, d.da, d.part);
}

__global__ void k_reduce_fixed_b(const DiagB* __restrict__ D, int nb)
{
    const DiagB& d = D[blockIdx.x];
    reduce_fixed_body(d.part, nb, 8, d.sums);
}

__global__ void k_spectra_b(const int* __restrict__ modes, const int* __restrict__ offs, const DiagB* __restrict__ D)
{
    const DiagB& d = D[blockIdx.y];
    spectra_body(blockIdx.x, modes, offs, d.KW1, d.KW2, d.KQ1, d.KQ2, d.KR, d.KP, d.sa, d.spec);
}

// ----------------------------------------------------------------------------
// Particles (src/gpe_particle.f90)
// ----------------------------------------------------------------------------

// V(x,y) = V0 exp(-r^2/(2 d^2)), r from the periodic coordinate
// (1/c)(s + s^3/6 + s^5/40), s = sin(c x), c = pi/L  (:34-56)
__global__ void k_pot_x(cplx* __restrict__ X, int N, double dx, double c1, double V0, double d)
{
    const size_t n2 = (size_t)N * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        const int ix = (int)(i % N), iy = (int)(i / N);
        const double x = ix * dx, y = iy * dx;
        const double sx = sin(c1 * x), sy = sin(c1 * y);
        const double sx2 = sx * sx, sy2 = sy * sy;
        const double xp = (1.0 / c1) * (sx + (1.0 / 6.0) * (sx2 * sx) + (1.0 / 40.0) * (sx2 * sx2 * sx));
        const double yp = (1.0 / c1) * (sy + (1.0 / 6.0) * (sy2 * sy) + (1.0 / 40.0) * (sy2 * sy2 * sy));
        const double rp = sqrt(xp * xp + yp * yp);
        X[i] = mk(V0 * exp(-(rp * rp) / (2.0 * (d * d))), 0.0);
    }
}

// EX[p*N + i] = exp(-i kx(i) qx_p), EY[p*N + i] = exp(-i ky(i) qy_p)
__global__ void k_ptab(const double* __restrict__ pos, cplx* __restrict__ EX, cplx* __restrict__ EY,
                       int np, int N)
{
    const int n = np * N;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int pp = i / N, k = i % N;
        const double kk = (double)ksg(k, N);
        double sx, cx, sy, cy;
        sincos(kk * pos[2 * pp], &sx, &cx);
        sincos(kk * pos[2 * pp + 1], &sy, &cy);
        EX[i] = mk(cx, -sx);
        EY[i] = mk(cy, -sy);
    }
}

// KV(kx,ky) = Vhat(kx,ky) * sum_p EX_p(kx) EY_p(ky) on the unmasked modes of
// rows kx in [0, nrows) (nrows = N: all rows)   (:334-349, separable)
__global__ void k_kv(const cplx* __restrict__ Vhat, const cplx* __restrict__ EX, const cplx* __restrict__ EY,
                     cplx* __restrict__ KV, int np, int N, int rows_pos)
{
    extern __shared__ double dsh[];
    cplx* ex = reinterpret_cast<cplx*>(dsh);
    const int kx = blockIdx.x;
    for (int pp = threadIdx.x; pp < np; pp += blockDim.x) ex[pp] = EX[(size_t)pp * N + kx];
    __syncthreads();
    const int kxs = ksg(kx, N);
    (void)rows_pos;
    for (int ky = threadIdx.x; ky < N; ky += blockDim.x) {
        const int kys = ksg(ky, N);
        if (kmasked(kxs, kys, N)) continue;
        cplx sacc = czero();
        for (int pp = 0; pp < np; ++pp) sacc = cadd(sacc, cmul(ex[pp], EY[(size_t)pp * N + ky]));
        const size_t o = (size_t)kx * N + ky;
        KV[o] = cmul(sacc, Vhat[o]);
    }
}

// Force partial sums for rows kx: for each particle p (PG particles per
// block in blockIdx.y), B_p = sum_ky A conj(EY_p(ky)), By_p = sum_ky ky A
// conj(EY_p(ky)) with A = rhohat_d(k) * conj(Vhat(k)) = rhohat_d(k) Vhat(-k)
// (V real).  out: fpart[((kx*np)+p)*4 + {ReB, ImB, ReBy, ImBy}]
template <int PG>
__global__ void k_force_rows(const cplx* __restrict__ Krho, const cplx* __restrict__ Vhat,
                             const cplx* __restrict__ EY, double* __restrict__ fpart, int np, int N)
{
    __shared__ double sh[4 * PG * 4];
    const int kx = blockIdx.x, p0 = blockIdx.y * PG;
    const int kxs = ksg(kx, N);
    double acc[4 * PG];
#pragma unroll
    for (int i = 0; i < 4 * PG; ++i) acc[i] = 0.0;
    for (int ky = threadIdx.x; ky < N; ky += blockDim.x) {
        const int kys = ksg(ky, N);
        if (kmasked(kxs, kys, N)) continue;
        const size_t o = (size_t)kx * N + ky;
        const cplx A = cmul(Krho[o], conjc(Vhat[o]));
#pragma unroll
        for (int j = 0; j < PG; ++j) {
            const int pp = p0 + j;
            if (pp < np) {
                const cplx z = cmul(A, conjc(EY[(size_t)pp * N + ky]));
                acc[4 * j + 0] += z.x;
                acc[4 * j + 1] += z.y;
                acc[4 * j + 2] += (double)kys * z.x;
                acc[4 * j + 3] += (double)kys * z.y;
            }
        }
    }
    block_sum<4 * PG>(acc, sh);
    if (threadIdx.x == 0)
        for (int j = 0; j < PG; ++j)
            if (p0 + j < np)
                for (int k = 0; k < 4; ++k) fpart[((size_t)kx * np + p0 + j) * 4 + k] = acc[4 * j + k];
}

// fl[p] = coef * sum over the row blocks b whose window flag is set, in a
// fixed order, of the nw wave partials ppart[b][w][p] (x, y); one workgroup
// per particle
__global__ void k_pforce_rs(const double* __restrict__ ppart, const unsigned char* __restrict__ phit, int nblk,
                            int nw, int np, double coef, double* __restrict__ fl)
{
    __shared__ double sh[2 * 16];
    const int pp = blockIdx.x;
    double acc[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        if (!phit[(size_t)b * np + pp]) continue;
        for (int w = 0; w < nw; ++w) {
            acc[0] += ppart[(((size_t)b * nw + w) * np + pp) * 2];
            acc[1] += ppart[(((size_t)b * nw + w) * np + pp) * 2 + 1];
        }
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) {
        fl[2 * pp] = coef * acc[0];
        fl[2 * pp + 1] = coef * acc[1];
    }
}

struct PartArgs {
    int src;                  // 0: spectral row partials (fpart); 1: real-space block partials
    int nblk;                 // real-space: number of row blocks
    double coef_rs;           // real-space: 2 alpha dvol / N^2
    int np, N, nrows, half;   // rows kx in [0,nrows); half: weight 2 for kx>0
    double coef;              // 2 alpha dvol / N^2
    double rf;                // short-range factor 384 dE pi^12 r0^12 / L^12
    double dt, mass;
    int stage;                // 1..4 RK4 stage; 0: forces only (diagnostics)
    double efac;              // 64 dE pi^12 r0^12 / L^12 (diagnostic energy)
};

// Particle forces and one RK4 stage of gp_particle_motion_rk4
// (src/gpe_particle.f90:219-324, 378-531), one workgroup of PSTEP_T threads.
// Fluid force F_j = -2 alpha dvol Re sum_x d_j rho V_p
//   = (2 alpha dvol/N^2) Im sum_k k_j rhohat_d(k) Vhat(-k) e^{i k.q_p}.
// The reference advances the particles one at a time, so the short-range
// force on particle ip sees the particles q < ip at their new positions and
// q > ip at their old ones.  A particle's new position depends only on the
// state at the start of the stage (tpos, its velocity, the earlier stages'
// kq), so all new positions are formed first; the forces then run one
// particle per wave, each with the lane order and shuffle tree of the
// one-wave form (lane l sums q = l, l + 64, ...), from positions in LDS.
// st: pos(2np) vel(2np) fext(2np) tpos tvel ku[4](2np) kq[4](2np) frc(2np)
//     then diag[5] = {eobj_kin, esrep, momx, momy, -}
constexpr int PSTEP_T = 512;
__global__ void __launch_bounds__(PSTEP_T) k_pstep(const double* __restrict__ fpart, const cplx* __restrict__ EX,
                                                   double* __restrict__ st, PartArgs a)
{
    __shared__ double fl[2 * 256];
    __shared__ double po[2 * 256];   // positions at the start of the stage
    __shared__ double pn[2 * 256];   // positions after the stage
    __shared__ double ff[2 * 256];   // total force
    const int np = a.np;
    double* pos = st;
    double* vel = st + 2 * np;
    double* fext = st + 4 * np;
    double* tpos = st + 6 * np;
    double* tvel = st + 8 * np;
    double* ku = st + 10 * np;   // [4][2np]
    double* kq = st + 18 * np;   // [4][2np]
    double* frc = st + 26 * np;
    double* dg = st + 28 * np;
    const int tid = threadIdx.x, lane = tid % 64, wv = tid / 64;
    constexpr int NW = PSTEP_T / 64;
    // fluid forces, fixed order over rows (spectral) or row blocks (real space)
    for (int pp = tid; pp < np && a.src == 1; pp += PSTEP_T) {
        // reduced by k_pforce_rs into st[28 np + 8 ...]
        fl[2 * pp] = st[28 * np + 8 + 2 * pp];
        fl[2 * pp + 1] = st[28 * np + 8 + 2 * pp + 1];
    }
    for (int pp = tid; pp < np && a.src == 0; pp += PSTEP_T) {
        double bx = 0.0, by = 0.0;
        for (int kx = 0; kx < a.nrows; ++kx) {
            const double* f = fpart + ((size_t)kx * np + pp) * 4;
            const double w = (a.half && kx > 0) ? 2.0 : 1.0;
            const cplx e = conjc(EX[(size_t)pp * a.N + kx]);   // e^{+i kx qx}
            const double kxd = (double)ksg(kx, a.N);
            const cplx B = cmul(e, mk(f[0], f[1]));
            const cplx By = cmul(e, mk(f[2], f[3]));
            bx += w * kxd * B.y;
            by += w * By.y;
        }
        fl[2 * pp] = a.coef * bx;
        fl[2 * pp + 1] = a.coef * by;
    }
    // start-of-stage positions, and the new positions
    // (stage 1 first saves tpos, tvel)
    for (int i = tid; i < 2 * np; i += PSTEP_T) {
        const double q = pos[i], u = vel[i];
        if (a.stage == 1) { tpos[i] = q; tvel[i] = u; }
        po[i] = q;
        double qn = q;
        if (a.stage > 0) {
            const double kqv = a.dt * u;
            const double tp = a.stage == 1 ? q : tpos[i];
            if (a.stage <= 2) qn = tp + 0.5 * kqv;
            else if (a.stage == 3) qn = tp + kqv;
            else qn = tp + (1.0 / 6.0) * (kq[i] + 2.0 * kq[2 * np + i] + 2.0 * kq[4 * np + i] + kqv);
        }
        pn[i] = qn;
    }
    __syncthreads();
    // short-range forces, one particle per wave
    for (int ip = wv; ip < np; ip += NW) {
        double sx = 0.0, sy = 0.0;
        const double px = po[2 * ip], py = po[2 * ip + 1];
        for (int qq = lane; qq < np; qq += 64) {
            if (qq == ip) continue;
            const double* P = (a.stage > 0 && qq < ip) ? pn : po;
            const double x = px - P[2 * qq], y = py - P[2 * qq + 1];
            const double b = -3.0 + cos(x) + cos(y);
            const double b2 = b * b, b4 = b2 * b2;
            const double den = b4 * b2 * b;
            sx += -a.rf * sin(x) / den;
            sy += -a.rf * sin(y) / den;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
        }
        if (np == 1) { sx = 0.0; sy = 0.0; }
        if (lane == 0) {
            ff[2 * ip] = fl[2 * ip] + sx + fext[2 * ip];
            ff[2 * ip + 1] = fl[2 * ip + 1] + sy + fext[2 * ip + 1];
        }
    }
    __syncthreads();
    // stage update of every particle
    for (int i = tid; i < 2 * np; i += PSTEP_T) {
        const double F = ff[i];
        frc[i] = F;
        if (a.stage > 0) {
            const int sidx = a.stage - 1;
            const double kuv = a.dt * F / a.mass;
            const double u = vel[i];
            ku[sidx * 2 * np + i] = kuv;
            kq[sidx * 2 * np + i] = a.dt * u;
            const double tv = a.stage == 1 ? u : tvel[i];
            if (a.stage <= 2) vel[i] = tv + 0.5 * kuv;
            else if (a.stage == 3) vel[i] = tv + kuv;
            else vel[i] = tv + (1.0 / 6.0) * (ku[i] + 2.0 * ku[2 * np + i] + 2.0 * ku[4 * np + i] + kuv);
            pos[i] = pn[i];
        }
    }
    if (a.stage == 0 && wv == 0) {
        // particle energies and momentum (src/diagnostics.f90:210-239)
        double ek = 0.0, es = 0.0, mx = 0.0, my = 0.0;
        for (int i1 = lane; i1 < np; i1 += 64) {
            ek += vel[2 * i1] * vel[2 * i1] + vel[2 * i1 + 1] * vel[2 * i1 + 1];
            mx += a.mass * vel[2 * i1];
            my += a.mass * vel[2 * i1 + 1];
            for (int i2 = 0; i2 < np; ++i2) {
                if (i1 == i2) continue;
                const double x = po[2 * i1] - po[2 * i2], y = po[2 * i1 + 1] - po[2 * i2 + 1];
                const double b = -3.0 + cos(x) + cos(y);
                const double b2 = b * b;
                es += a.efac / (b2 * b2 * b2);
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            ek += __shfl_xor(ek, o, 64);
            es += __shfl_xor(es, o, 64);
            mx += __shfl_xor(mx, o, 64);
            my += __shfl_xor(my, o, 64);
        }
        if (lane == 0) {
            dg[0] = 0.5 * a.mass * ek;
            dg[1] = np == 1 ? 0.0 : 0.5 * es;
            dg[2] = mx;
            dg[3] = my;
        }
    }
}

// ----------------------------------------------------------------------------
// SGLE (src/gpe.f90:149-217, 226-266)
// ----------------------------------------------------------------------------

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32x32->64 multiplications
// and a Weyl-sequence key schedule; counter c, key k.
__host__ __device__ inline void philox4x32_10(unsigned c[4], unsigned k0, unsigned k1)
{
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = 0xD2511F53ull * c[0], p1 = 0xCD9E8D57ull * c[2];
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Two independent standard normals at grid point i for noise draw `step`
// (counter (i, 0, step_lo, step_hi), key = seed): uniforms from 53-bit
// fractions of the 64-bit words, u1 in (0, 1], u2 in [0, 1), then the
// Box-Muller transform.  The reference draws its pair with the polar method
// from Fortran's RANDOM_NUMBER; only the distribution is the same.
__host__ __device__ inline void sgle_normals(unsigned long long i, unsigned long long step, unsigned long long seed,
                                             double* v1, double* v2)
{
    unsigned c[4] = {(unsigned)i, (unsigned)(i >> 32), (unsigned)step, (unsigned)(step >> 32)};
    philox4x32_10(c, (unsigned)seed, (unsigned)(seed >> 32));
    const unsigned long long a = ((unsigned long long)c[1] << 32) | c[0];
    const unsigned long long b = ((unsigned long long)c[3] << 32) | c[2];
    const double u1 = (double)((a >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(b >> 11) * 0x1.0p-53;
    const double rr = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586476925 * u2, &sn, &cs);
    *v1 = rr * cs;
    *v2 = rr * sn;
}

// X = noise field of draw `step`, times (1/sqrt(dvol)) then noisesigma
// (src/gpe.f90:163), or the caller's standard normals in X scaled in place
// (gen = 0)
__global__ void k_sgle_noise(cplx* __restrict__ X, int N, unsigned long long step, unsigned long long seed,
                             double invsq, double sigma, int gen)
{
    const size_t n2 = (size_t)N * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        double v1, v2;
        if (gen) sgle_normals(i, step, seed, &v1, &v2);
        else {
            v1 = X[i].x;
            v2 = X[i].y;
        }
        X[i] = rmul(sigma, rmul(invsq, mk(v1, v2)));
    }
}

// block partial sums of |K|^2 over a k-space field
__global__ void k_knorm(const cplx* __restrict__ K, size_t n2, double* __restrict__ part)
{
    __shared__ double sh[16];
    double acc[1] = {0.0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        const cplx z = K[i];
        acc[0] += z.x * z.x + z.y * z.y;
    }
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// finiteT_adhoc_chempot (src/gpe.f90:255-264): norm = sum_x |psi|^2 dvol =
// (dvol/N^2) sum_k |psi-hat|^2 (Parseval), omega -= dt (nuN/L^2)(norm - L^2)
__global__ void k_chempot(const double* __restrict__ part, int nb, double* __restrict__ omega, double dvol_n2,
                          double dt, double nuN, double L2)
{
    __shared__ double sh[16];
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x) acc[0] += part[b];
    block_sum<1>(acc, sh);
    if (threadIdx.x == 0) {
        const double norm = acc[0] * dvol_n2;
        omega[0] = omega[0] - dt * (nuN / L2) * (norm - L2);
    }
}

// ----------------------------------------------------------------------------
// Vortex initial condition and the imposed vortex velocity
// ----------------------------------------------------------------------------

// Padé-core vortex product: amplitude prod_j sqrt(rho(r_j/xi)), r_j the
// minimum-image distance; phase sum_j q_j sum_{|m|<=nimg}
// arg(sin((x-x_j)/2 + i(y-y_j-2 pi m)/2)) - (sum_j q_j x_j) y/(2 pi), which
// is 2 pi-periodic in x and y for zero net charge up to exp(-2 pi nimg).
__host__ __device__ inline void ic_point(double x, double y, int nv, const double* xs, const double* ys,
                                         const int* qs, double xi, int nimg, double* out)
{
    const double pi = 3.14159265358979323846;
    const double a1 = 11.0 / 32.0, a2 = 11.0 / 384.0, b1 = 1.0 / 3.0;
    double amp = 1.0, pr = 1.0, pim = 0.0, lin = 0.0;
    for (int j = 0; j < nv; ++j) {
        double dxp = x - xs[j], dyp = y - ys[j];
        double wx = dxp - 2.0 * pi * floor(dxp / (2.0 * pi) + 0.5);
        double wy = dyp - 2.0 * pi * floor(dyp / (2.0 * pi) + 0.5);
        double r2 = (wx * wx + wy * wy) / (xi * xi);
        double rho = r2 * (a1 + a2 * r2) / (1.0 + b1 * r2 + a2 * r2 * r2);
        amp *= sqrt(rho);
        lin += qs[j] * xs[j];
        const double sX = sin(0.5 * dxp), cX = cos(0.5 * dxp);
        for (int m = -nimg; m <= nimg; ++m) {
            const double Y = 0.5 * (dyp - 2.0 * pi * m);
            double zr = sX * cosh(Y), zi = cX * sinh(Y);
            const double h = sqrt(zr * zr + zi * zi);
            if (h > 0.0) { zr /= h; zi /= h; } else { zr = 1.0; zi = 0.0; }
            if (qs[j] < 0) zi = -zi;
            const double nr = pr * zr - pim * zi, ni = pr * zi + pim * zr;
            pr = nr;
            pim = ni;
        }
        const double hn = sqrt(pr * pr + pim * pim);
        pr /= hn;
        pim /= hn;
    }
    const double th = -lin * y / (2.0 * pi);
    const double c = cos(th), s = sin(th);
    out[0] = amp * (pr * c - pim * s);
    out[1] = amp * (pr * s + pim * c);
}

// Imposed velocity of a vortex configuration for the advective ARGLE:
// u = 2 alpha rho0 grad(theta0), the current of the Padé product ansatz
// (rho0 = prod_j rho(r_j/xi), theta0 its phase), smooth and zero in the
// cores.  grad arg(sin((X + i Y_m)/2)) = (-sinh Y_m, sin X)/(2 (cosh Y_m -
// cos X)), summed over the nimg images Y_m = Y - 2 pi m and the vortices
// with their charges, minus (sum_j q_j x_j)/(2 pi) along y from the linear
// phase term.
__host__ __device__ inline void vortex_velocity_point(double x, double y, int nv, const double* xs,
                                                      const double* ys, const int* qs, double xi, int nimg,
                                                      double alpha, double* ux, double* uy)
{
    const double pi = 3.14159265358979323846;
    const double a1 = 11.0 / 32.0, a2 = 11.0 / 384.0, b1 = 1.0 / 3.0;
    double rho = 1.0, gx = 0.0, gy = 0.0, lin = 0.0;
    for (int j = 0; j < nv; ++j) {
        const double dxp = x - xs[j], dyp = y - ys[j];
        const double wx = dxp - 2.0 * pi * floor(dxp / (2.0 * pi) + 0.5);
        const double wy = dyp - 2.0 * pi * floor(dyp / (2.0 * pi) + 0.5);
        const double r2 = (wx * wx + wy * wy) / (xi * xi);
        rho *= r2 * (a1 + a2 * r2) / (1.0 + b1 * r2 + a2 * r2 * r2);
        lin += qs[j] * xs[j];
        const double sX = sin(dxp), cX = cos(dxp);
        for (int m = -nimg; m <= nimg; ++m) {
            const double Ym = dyp - 2.0 * pi * m;
            const double den = cosh(Ym) - cX;
            if (den > 0.0) {
                gx += qs[j] * (-0.5 * sinh(Ym) / den);
                gy += qs[j] * (0.5 * sX / den);
            }
        }
    }
    gy -= lin / (2.0 * pi);
    *ux = 2.0 * alpha * rho * gx;
    *uy = 2.0 * alpha * rho * gy;
}

__global__ void k_vortex_velocity(double* __restrict__ UX, double* __restrict__ UY, int N, int nv,
                                  const double* __restrict__ xs, const double* __restrict__ ys,
                                  const int* __restrict__ qs, double xi, int nimg, double alpha)
{
    const size_t n2 = (size_t)N * N;
    const double dx = 2.0 * 3.14159265358979323846 / N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        vortex_velocity_point((int)(i % N) * dx, (int)(i / N) * dx, nv, xs, ys, qs, xi, nimg, alpha, UX + i, UY + i);
}

__global__ void k_ic(cplx* __restrict__ X, int N, int nv, const double* __restrict__ xs,
                     const double* __restrict__ ys, const int* __restrict__ qs, double xi, int nimg)
{
    const size_t n2 = (size_t)N * N;
    const double dx = 2.0 * 3.14159265358979323846 / N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
        const int ix = (int)(i % N), iy = (int)(i / N);
        double o[2];
        ic_point(ix * dx, iy * dx, nv, xs, ys, qs, xi, nimg, o);
        X[i] = mk(o[0], o[1]);
    }
}

}  // namespace vk
