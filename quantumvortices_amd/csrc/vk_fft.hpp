// Register-resident Stockham FFT for CDNA4 (gfx950), FP64 complex.
//
// One length-N transform is held by T = N/S threads; thread t owns the S
// points at positions t + T*s (s = 0..S-1), both on input and on output.
// Stages are radix R (R | S, largest first); a stage with sub-transform
// length Ns reads the butterfly j = t + T*a from registers v[a + (S/R)*r],
// applies twiddles w_{Ns*R}^{r*(j%Ns)}, does the radix-R DFT and writes to
// position (j/Ns)*Ns*R + j%Ns + r*Ns.  Between stages the points go through
// LDS (padded slots p + p/16, lslot below, against bank conflicts of the
// strided writes); the last stage's outputs land back in the same
// register slots, so a global load feeds the first stage directly and the
// last stage feeds a global store directly.  N = R^k needs k-1 exchanges.
//
// Only the forward (sign -1) transform is written; the inverse is
// conj(FFT(conj(x))), folded into the callers' loads and stores.
#pragma once
#include <hip/hip_runtime.h>

namespace vk {

// Workgroup barrier that orders LDS only.  The transforms exchange data
// through LDS alone, so their barriers need not wait for this thread's
// outstanding global loads and stores (a plain __syncthreads() waits for
// all of them, which serialises stores, and asynchronous global->LDS
// copies, into every exchange).
#ifndef VK_LDS_BARRIER
#define VK_LDS_BARRIER 1
#endif
__device__ __forceinline__ void lds_barrier()
{
#if VK_LDS_BARRIER
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
#else
    __syncthreads();
#endif
}

typedef double2 cplx;

__device__ __forceinline__ cplx mk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cplx conjc(cplx a) { return mk(a.x, -a.y); }
__device__ __forceinline__ cplx rmul(double r, cplx a) { return mk(r * a.x, r * a.y); }
// a * (-i)
__device__ __forceinline__ cplx mul_mi(cplx a) { return mk(a.y, -a.x); }

// cos(pi/8), sin(pi/8), sqrt(1/2)
constexpr double kC8 = 0.92387953251128675613;
constexpr double kS8 = 0.38268343236508977173;
constexpr double kR2 = 0.70710678118654752440;

__device__ __forceinline__ void dft2(cplx& a, cplx& b)
{
    cplx t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// in: x0..x3, out natural order X0..X3 (forward)
__device__ __forceinline__ void dft4(cplx& x0, cplx& x1, cplx& x2, cplx& x3)
{
    cplx s02 = cadd(x0, x2), d02 = csub(x0, x2);
    cplx s13 = cadd(x1, x3), d13 = mul_mi(csub(x1, x3));
    x0 = cadd(s02, s13);
    x2 = csub(s02, s13);
    x1 = cadd(d02, d13);
    x3 = csub(d02, d13);
}

// multiply by w_8^k (forward), k = 0..3
template <int k> __device__ __forceinline__ cplx w8(cplx a);
template <> __device__ __forceinline__ cplx w8<0>(cplx a) { return a; }
template <> __device__ __forceinline__ cplx w8<1>(cplx a)
{
    return mk(kR2 * (a.x + a.y), kR2 * (a.y - a.x));
}
template <> __device__ __forceinline__ cplx w8<2>(cplx a) { return mul_mi(a); }
template <> __device__ __forceinline__ cplx w8<3>(cplx a)
{
    return mk(kR2 * (a.y - a.x), -kR2 * (a.x + a.y));
}

// multiply by w_16^k (forward) for the exponents a 4x4 split needs
template <int k> __device__ __forceinline__ cplx w16(cplx a)
{
    // generic: angle -2 pi k / 16
    constexpr int m = k % 16;
    if constexpr (m == 0) return a;
    else if constexpr (m == 1) return mk(kC8 * a.x + kS8 * a.y, kC8 * a.y - kS8 * a.x);
    else if constexpr (m == 2) return w8<1>(a);
    else if constexpr (m == 3) return mk(kS8 * a.x + kC8 * a.y, kS8 * a.y - kC8 * a.x);
    else if constexpr (m == 4) return mul_mi(a);
    else if constexpr (m == 6) return w8<3>(a);
    else if constexpr (m == 9) return mk(-kC8 * a.x - kS8 * a.y, -kC8 * a.y + kS8 * a.x);
    else { static_assert(m == 0 || m == 1, "unused w16 exponent"); return a; }
}

template <int R> struct Dft;

template <> struct Dft<2> {
    __device__ __forceinline__ static void run(cplx* b) { dft2(b[0], b[1]); }
};
template <> struct Dft<4> {
    __device__ __forceinline__ static void run(cplx* b) { dft4(b[0], b[1], b[2], b[3]); }
};
// 8 = 2 x 4 (decimation in frequency)
template <> struct Dft<8> {
    __device__ __forceinline__ static void run(cplx* b)
    {
        cplx e0 = cadd(b[0], b[4]), e1 = cadd(b[1], b[5]);
        cplx e2 = cadd(b[2], b[6]), e3 = cadd(b[3], b[7]);
        cplx o0 = csub(b[0], b[4]);
        cplx o1 = w8<1>(csub(b[1], b[5]));
        cplx o2 = w8<2>(csub(b[2], b[6]));
        cplx o3 = w8<3>(csub(b[3], b[7]));
        dft4(e0, e1, e2, e3);
        dft4(o0, o1, o2, o3);
        b[0] = e0; b[2] = e1; b[4] = e2; b[6] = e3;
        b[1] = o0; b[3] = o1; b[5] = o2; b[7] = o3;
    }
};
// 16 = 4 x 4: X[k2 + 4 k1] = sum_j w4^{j k1} w16^{j k2} sum_m a[j+4m] w4^{m k2}
template <> struct Dft<16> {
    __device__ __forceinline__ static void run(cplx* b)
    {
        cplx y[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cplx a0 = b[j], a1 = b[j + 4], a2 = b[j + 8], a3 = b[j + 12];
            dft4(a0, a1, a2, a3);
            y[j][0] = a0; y[j][1] = a1; y[j][2] = a2; y[j][3] = a3;
        }
        y[1][1] = w16<1>(y[1][1]); y[1][2] = w16<2>(y[1][2]); y[1][3] = w16<3>(y[1][3]);
        y[2][1] = w16<2>(y[2][1]); y[2][2] = w16<4>(y[2][2]); y[2][3] = w16<6>(y[2][3]);
        y[3][1] = w16<3>(y[3][1]); y[3][2] = w16<6>(y[3][2]); y[3][3] = w16<9>(y[3][3]);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) {
            cplx a0 = y[0][k2], a1 = y[1][k2], a2 = y[2][k2], a3 = y[3][k2];
            dft4(a0, a1, a2, a3);
            b[k2] = a0; b[k2 + 4] = a1; b[k2 + 8] = a2; b[k2 + 12] = a3;
        }
    }
};

// Radix plan: largest radix dividing both S and the remaining length.
__host__ __device__ constexpr int radix_at(int N, int S, int st)
{
    int n = N;
    int r = 1;
    for (int i = 0; i <= st; ++i) {
        r = 1;
        for (int R = 16; R >= 2; R >>= 1)
            if (S % R == 0 && n % R == 0) { r = R; break; }
        n /= r;
    }
    return r;
}

// LDS slot of point p in an exchange: one pad double per 16 points, so that
// the 16-strided writes of a stage with sub-length 1 spread over the banks
__device__ __forceinline__ int lslot(int p) { return p + (p >> 4); }

// LDS words per transform for one real component
template <int N> struct Lds {
    static constexpr int words = N + N / 16;
};

// lslot(t + T*s), written so that the compiler sees one base address and a
// constant offset per s where that is exact (16 | T): one address register
// for all S reads
template <int T> __device__ __forceinline__ int gather_idx(int t, int s)
{
    if constexpr (T % 16 == 0) return lslot(t) + s * (T + T / 16);
    else return lslot(t + T * s);
}

// lslot(idxD + r*NS) for the scatter of a stage with sub-length NS and
// radix R, again as base + constant offset where that is exact: NS a
// multiple of 16, or NS = 1 with R dividing 16 (idxD = R*j, and R*j + r
// crosses no multiple of 16 for r < R)
template <int NS, int R> __device__ __forceinline__ int scatter_idx(int idxD, int r)
{
    if constexpr (NS % 16 == 0) return lslot(idxD) + r * (NS + NS / 16);
    else if constexpr (NS == 1 && 16 % R == 0) return lslot(idxD) + r;
    else return lslot(idxD + r * NS);
}

// One Stockham stage, then recurse.  v: registers; re/im: this transform's
// LDS arrays (split storage: real parts and imaginary parts exchanged in
// turn through a single array of Lds<N>::words doubles when re == im).
template <int N, int S, int ST, int NS, bool SPLIT>
struct Stage {
    static constexpr int R = radix_at(N, S, ST);
    static constexpr int T = N / S;
    static constexpr bool LAST = (NS * R == N);

    __device__ __forceinline__ static void run(cplx (&v)[S], double* re, double* im, int t,
                                               const cplx* __restrict__ tw)
    {
        static_assert(S % R == 0, "radix must divide S");
#pragma unroll
        for (int a = 0; a < S / R; ++a) {
            const int j = t + T * a;
            const int jm = j % NS;
            cplx x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = v[a + (S / R) * r];
            if constexpr (NS > 1) {
                // w^r for r = 4a + b from table values w^1..w^3 and w^4a
                // (one product at most per factor)
                constexpr int STEP = N / (NS * R);
                const int base = jm * STEP;
                const cplx w1 = tw[base];
                if constexpr (R == 2) {
                    x[1] = cmul(x[1], w1);
                } else {
                    const cplx w2 = tw[2 * base], w3 = tw[3 * base];
                    x[1] = cmul(x[1], w1);
                    x[2] = cmul(x[2], w2);
                    x[3] = cmul(x[3], w3);
#pragma unroll
                    for (int a4 = 1; a4 < R / 4; ++a4) {
                        const cplx wa = tw[4 * a4 * base];
                        x[4 * a4] = cmul(x[4 * a4], wa);
                        x[4 * a4 + 1] = cmul(x[4 * a4 + 1], cmul(wa, w1));
                        x[4 * a4 + 2] = cmul(x[4 * a4 + 2], cmul(wa, w2));
                        x[4 * a4 + 3] = cmul(x[4 * a4 + 3], cmul(wa, w3));
                    }
                }
            }
            Dft<R>::run(x);
#pragma unroll
            for (int r = 0; r < R; ++r) v[a + (S / R) * r] = x[r];
        }
        if constexpr (!LAST) {
            // scatter to LDS at Stockham positions, gather back t + T*s
            if constexpr (SPLIT) {
#pragma unroll
                for (int a = 0; a < S / R; ++a) {
                    const int j = t + T * a;
                    const int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
                    for (int r = 0; r < R; ++r) re[scatter_idx<NS, R>(idxD, r)] = v[a + (S / R) * r].x;
                }
                lds_barrier();
#pragma unroll
                for (int s = 0; s < S; ++s) v[s].x = re[gather_idx<T>(t, s)];
                lds_barrier();
#pragma unroll
                for (int a = 0; a < S / R; ++a) {
                    const int j = t + T * a;
                    const int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
                    for (int r = 0; r < R; ++r) re[scatter_idx<NS, R>(idxD, r)] = v[a + (S / R) * r].y;
                }
                lds_barrier();
#pragma unroll
                for (int s = 0; s < S; ++s) v[s].y = re[gather_idx<T>(t, s)];
                lds_barrier();
            } else {
#pragma unroll
                for (int a = 0; a < S / R; ++a) {
                    const int j = t + T * a;
                    const int idxD = (j / NS) * NS * R + (j % NS);
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        re[scatter_idx<NS, R>(idxD, r)] = v[a + (S / R) * r].x;
                        im[scatter_idx<NS, R>(idxD, r)] = v[a + (S / R) * r].y;
                    }
                }
                lds_barrier();
#pragma unroll
                for (int s = 0; s < S; ++s) v[s] = mk(re[gather_idx<T>(t, s)], im[gather_idx<T>(t, s)]);
                lds_barrier();
            }
            Stage<N, S, ST + 1, NS * R, SPLIT>::run(v, re, im, t, tw);
        }
    }
};

// Terminator for unused recursion depth (never instantiated at run time).
template <int N, int S, int ST, bool SPLIT> struct Stage<N, S, ST, N, SPLIT> {
    __device__ __forceinline__ static void run(cplx (&)[S], double*, double*, int, const cplx*) {}
};

// Forward transform of the points held by this thread group.
// lds: base of this transform's LDS region (Lds<N>::words doubles when
// SPLIT, twice that otherwise).  All threads of the workgroup must call it
// (it contains workgroup barriers).
template <int N, int S, bool SPLIT>
__device__ __forceinline__ void fft_fwd(cplx (&v)[S], double* lds, int t, const cplx* __restrict__ tw)
{
    // Register-only compiler barrier on the (uniform) table pointer: each
    // transform reloads its twiddles instead of keeping the previous
    // transform's values live across the pointwise work in between.
    const cplx* twb = tw;
    asm volatile("" : "+s"(twb));
    Stage<N, S, 0, 1, SPLIT>::run(v, lds, SPLIT ? lds : lds + Lds<N>::words, t, twb);
}

}  // namespace vk
