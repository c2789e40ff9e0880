// Fourier neural operator for the ARGLE warm start: forward, backward and
// Adam on the GPU.  Activations and accumulation are FP32, weights are BF16
// (FP32 master copies are kept for the optimizer).  No reference counterpart:
// the reference prepares ARGLE states from an analytic ansatz only
// (src/initial_config.f90, restart_from_argle_data).
//
// Network (per sample, N x N grid, planes [b][c][y][x], x fastest):
//   z0 = Wp x + bp                                    (lift, cin -> C)
//   u_{l+1} = W_l z_l + b_l + K_l z_l,  z_{l+1} = gelu(u_{l+1}) for l < L-1,
//   z_L = u_L
//   h = Q1 z_L + bq1 (C -> P),  out = Q2 gelu(h) + bq2 (P -> 2)
// Spectral convolution K z = E(R . D z):
//   D z (kx, ky) = sum_{x,y} z e^{-2 pi i (kx x + ky y)/N},  kx in [0,M),
//                  ky in [-M, M)  (index j = ky mod 2M)
//   (R . X)_o = s(kx) sum_i X_i R_io,  s(0) = 1/N^2, s(kx>0) = 2/N^2
//   E(Y)(x, y) = Re sum_{kx,ky} Y e^{+2 pi i (kx x + ky y)/N}
// Adjoints used by the backward pass (du = dL/du_{l+1}, real):
//   dR_io = s sum_b D(du_o) conj(X_i)
//   dz_i  = W^T du + Re sum_m [s sum_o D(du_o) conj(R_io)] e^{ik.x}
// Only 2 M^2 of the N^2 modes are kept, so D and E are separable sums over
// e^{-2 pi i k x/N} tables instead of full transforms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "vikshobha.h"

int vk_set_error(int code, const std::string& msg);   // vk_gpe.hip (vk_last_error)

namespace {

#define FCHK(x)                                                                                            \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess) return vk_set_error(VK_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

typedef float2 c32;

__host__ __device__ inline float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest, ties to even (finite inputs)
    return (uint16_t)(u >> 16);
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

constexpr int CIN_MAX = 4;

// ---------------------------------------------------------------- kernels --

// z0[b][c][p] = sum_k Wp[c][k] x[b][k][p] + bp[c]
template <int C>
__global__ void k_lift(const float* __restrict__ x, const uint16_t* __restrict__ Wp, const uint16_t* __restrict__ bp,
                       float* __restrict__ z, int B, int cin, int NN)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)B * NN) return;
    const int b = (int)(i / NN), p = (int)(i % NN);
    float xi[CIN_MAX];
#pragma unroll
    for (int k = 0; k < CIN_MAX; ++k) xi[k] = k < cin ? x[((size_t)b * cin + k) * NN + p] : 0.f;
#pragma unroll 4
    for (int c = 0; c < C; ++c) {
        float s = bf2f(bp[c]);
#pragma unroll
        for (int k = 0; k < CIN_MAX; ++k)
            if (k < cin) s += bf2f(Wp[c * cin + k]) * xi[k];
        z[((size_t)b * C + c) * NN + p] = s;
    }
}

// A[q][y][kx] = sum_x z[q][y][x] F[x][kx]; block = 16 rows of plane q,
// thread (row r, kx); x streamed through LDS in chunks of 64
template <int M>
__global__ void k_dftx(const float* __restrict__ z, const c32* __restrict__ F, c32* __restrict__ A, int N)
{
    __shared__ float zs[16][65];
    __shared__ c32 fs[64][M];
    const int q = blockIdx.y, y0 = blockIdx.x * 16;
    const int r = threadIdx.x / M, kx = threadIdx.x % M;
    float ar = 0.f, ai = 0.f;
    for (int x0 = 0; x0 < N; x0 += 64) {
        for (int i = threadIdx.x; i < 16 * 64; i += blockDim.x)
            zs[i >> 6][i & 63] = z[((size_t)q * N + y0 + (i >> 6)) * N + x0 + (i & 63)];
        for (int i = threadIdx.x; i < 64 * M; i += blockDim.x) fs[i / M][i % M] = F[(size_t)(x0 + i / M) * M + i % M];
        __syncthreads();
#pragma unroll 8
        for (int x = 0; x < 64; ++x) {
            const float v = zs[r][x];
            ar += v * fs[x][kx].x;
            ai += v * fs[x][kx].y;
        }
        __syncthreads();
    }
    A[((size_t)q * N + y0 + r) * M + kx] = make_float2(ar, ai);
}

// X[q][kx][j] = sum_y A[q][y][kx] G[y][j]; one block per plane, thread (kx, j)
template <int M>
__global__ void k_dfty(const c32* __restrict__ A, const c32* __restrict__ G, c32* __restrict__ X, int N)
{
    const int q = blockIdx.x;
    const int kx = threadIdx.x / (2 * M), j = threadIdx.x % (2 * M);
    float xr = 0.f, xi = 0.f;
    const c32* Aq = A + (size_t)q * N * M;
    for (int y = 0; y < N; ++y) {
        const c32 a = Aq[y * M + kx], g = G[y * 2 * M + j];
        xr += a.x * g.x - a.y * g.y;
        xi += a.x * g.y + a.y * g.x;
    }
    X[((size_t)q * M + kx) * 2 * M + j] = make_float2(xr, xi);
}

// forward: Y[b][o][m] = s(m) sum_i X[b][i][m] R[i][o][m]
// adjoint: Y[b][i][m] = s(m) sum_o X[b][o][m] conj(R[i][o][m])
template <int C, int M>
__global__ void k_mix(const c32* __restrict__ X, const uint16_t* __restrict__ R, c32* __restrict__ Y, int B,
                      int adjoint, float inv_n2)
{
    constexpr int NM = 2 * M * M;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)B * C * NM) return;
    const int m = (int)(t % NM), o = (int)((t / NM) % C), b = (int)(t / ((size_t)NM * C));
    float yr = 0.f, yi = 0.f;
#pragma unroll 4
    for (int i = 0; i < C; ++i) {
        const size_t ridx = adjoint ? (((size_t)o * C + i) * NM + m) : (((size_t)i * C + o) * NM + m);
        const float rr = bf2f(R[2 * ridx]);
        const float ri = adjoint ? -bf2f(R[2 * ridx + 1]) : bf2f(R[2 * ridx + 1]);
        const c32 x = X[((size_t)b * C + i) * NM + m];
        yr += x.x * rr - x.y * ri;
        yi += x.x * ri + x.y * rr;
    }
    const float s = (m / (2 * M) == 0 ? 1.f : 2.f) * inv_n2;
    Y[((size_t)b * C + o) * NM + m] = make_float2(s * yr, s * yi);
}

// Bk[q][y][kx] = sum_j Y[q][kx][j] conj(G[y][j])
template <int M>
__global__ void k_idfty(const c32* __restrict__ Y, const c32* __restrict__ G, c32* __restrict__ Bk, int N, int nq)
{
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)nq * N * M) return;
    const int kx = (int)(t % M), y = (int)((t / M) % N), q = (int)(t / ((size_t)M * N));
    const c32* Yq = Y + ((size_t)q * M + kx) * 2 * M;
    const c32* Gy = G + (size_t)y * 2 * M;
    float br = 0.f, bi = 0.f;
#pragma unroll
    for (int j = 0; j < 2 * M; ++j) {
        const c32 a = Yq[j], g = Gy[j];
        br += a.x * g.x + a.y * g.y;
        bi += a.y * g.x - a.x * g.y;
    }
    Bk[t] = make_float2(br, bi);
}

// Pointwise channel mixing fused with the spectral inverse, one block per
// (y, b), threads over x:
//  trans = 0: s_c = sum_i W[c][i] zin_i + bias_c + sum_kx Re(Bk[c][kx] e^{2 pi i kx x/N})
//             u_c = s_c,  zout_c = act ? gelu(s_c) : s_c
//  trans = 1: s_i = sum_c W[c][i] zin_c + sum_kx Re(Bk[i][kx] e^{2 pi i kx x/N})
//             zout_i = uprev ? s_i gelu'(uprev_i) : s_i
template <int C, int M>
__global__ void k_out(const float* __restrict__ zin, const uint16_t* __restrict__ W, const uint16_t* __restrict__ bias,
                      const c32* __restrict__ Bk, const c32* __restrict__ F, float* __restrict__ u,
                      float* __restrict__ zout, const float* __restrict__ uprev, int N, int trans, int act)
{
    __shared__ float Ws[C * C];
    __shared__ c32 Bs[C * M];
    __shared__ float bs[C];
    const int b = blockIdx.y, y = blockIdx.x;
    for (int i = threadIdx.x; i < C * C; i += blockDim.x) {
        const int r = i / C, c = i % C;
        Ws[i] = trans ? bf2f(W[c * C + r]) : bf2f(W[i]);   // Ws[out][in]
    }
    for (int i = threadIdx.x; i < C * M; i += blockDim.x) {
        const int c = i / M, kx = i % M;
        Bs[i] = Bk[(((size_t)b * C + c) * N + y) * M + kx];
    }
    for (int i = threadIdx.x; i < C; i += blockDim.x) bs[i] = (!trans && bias) ? bf2f(bias[i]) : 0.f;
    __syncthreads();
    const size_t NN = (size_t)N * N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        float zi[C];
#pragma unroll
        for (int i = 0; i < C; ++i) zi[i] = zin[((size_t)b * C + i) * NN + (size_t)y * N + x];
        c32 f[M];
#pragma unroll
        for (int kx = 0; kx < M; ++kx) f[kx] = F[(size_t)x * M + kx];
#pragma unroll 2
        for (int c = 0; c < C; ++c) {
            float s = bs[c];
#pragma unroll
            for (int i = 0; i < C; ++i) s += Ws[c * C + i] * zi[i];
#pragma unroll
            for (int kx = 0; kx < M; ++kx) {
                const c32 a = Bs[c * M + kx];
                s += a.x * f[kx].x + a.y * f[kx].y;   // Re(a conj(f)), f = e^{-i...}
            }
            const size_t o = ((size_t)b * C + c) * NN + (size_t)y * N + x;
            if (!trans) {
                u[o] = s;
                zout[o] = act ? gelu(s) : s;
            } else {
                zout[o] = uprev ? s * gelu_d(uprev[o]) : s;
            }
        }
    }
}

// h = Q1 z + bq1 (stored), out = Q2 gelu(h) + bq2
template <int C>
__global__ void k_proj(const float* __restrict__ z, const uint16_t* __restrict__ Q1, const uint16_t* __restrict__ bq1,
                       const uint16_t* __restrict__ Q2, const uint16_t* __restrict__ bq2, float* __restrict__ h,
                       float* __restrict__ out, int B, int P, int NN)
{
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)B * NN) return;
    const int b = (int)(t / NN), p = (int)(t % NN);
    float zi[C];
#pragma unroll
    for (int i = 0; i < C; ++i) zi[i] = z[((size_t)b * C + i) * NN + p];
    float o0 = bf2f(bq2[0]), o1 = bf2f(bq2[1]);
    for (int j = 0; j < P; ++j) {
        float s = bf2f(bq1[j]);
#pragma unroll
        for (int i = 0; i < C; ++i) s += bf2f(Q1[j * C + i]) * zi[i];
        h[((size_t)b * P + j) * NN + p] = s;
        const float g = gelu(s);
        o0 += bf2f(Q2[j]) * g;
        o1 += bf2f(Q2[P + j]) * g;
    }
    out[((size_t)b * 2 + 0) * NN + p] = o0;
    out[((size_t)b * 2 + 1) * NN + p] = o1;
}

// dout = 2 (out - tgt) / n ; part[block] = sum (out - tgt)^2 / n
__global__ void k_loss(const float* __restrict__ out, const float* __restrict__ tgt, float* __restrict__ dout,
                       double* __restrict__ part, size_t n, float inv_n)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = out[i] - tgt[i];
        acc += (double)d * d;
        dout[i] = 2.f * d * inv_n;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] * inv_n;
}

// per pixel: dg = Q2^T dout, dh = dg gelu'(h), dz = Q1^T dh; stores dh and gelu(h)
template <int C>
__global__ void k_proj_bwd(const float* __restrict__ h, const float* __restrict__ dout,
                           const uint16_t* __restrict__ Q1, const uint16_t* __restrict__ Q2, float* __restrict__ dh,
                           float* __restrict__ gq, float* __restrict__ dz, int B, int P, int NN)
{
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)B * NN) return;
    const int b = (int)(t / NN), p = (int)(t % NN);
    const float d0 = dout[((size_t)b * 2) * NN + p], d1 = dout[((size_t)b * 2 + 1) * NN + p];
    float acc[C];
#pragma unroll
    for (int i = 0; i < C; ++i) acc[i] = 0.f;
    for (int j = 0; j < P; ++j) {
        const size_t o = ((size_t)b * P + j) * NN + p;
        const float hv = h[o];
        const float dhv = (bf2f(Q2[j]) * d0 + bf2f(Q2[P + j]) * d1) * gelu_d(hv);
        dh[o] = dhv;
        gq[o] = gelu(hv);
#pragma unroll
        for (int i = 0; i < C; ++i) acc[i] += bf2f(Q1[j * C + i]) * dhv;
    }
#pragma unroll
    for (int i = 0; i < C; ++i) dz[((size_t)b * C + i) * NN + p] = acc[i];
}

// Weight-gradient reduction.  part[blk][a][c] = sum over the block's pixel
// range (all samples) of A[b][a][p] * Bm[b][c][p]  (Bm == null: Bm = 1,
// nc = 1).  Tiles of 64 pixels go through LDS; each thread owns a 4 x 4
// block of outputs (na, nc <= 64).
__global__ void __launch_bounds__(256) k_outer(const float* __restrict__ A, const float* __restrict__ Bm,
                                               double* __restrict__ part, int B, int na, int nc, int NN, int chunk)
{
    __shared__ float sa[64][65];
    __shared__ float sb[64][65];
    const int ncc = Bm ? nc : 1;
    const int ta = (na + 3) / 4, tc = (ncc + 3) / 4;
    const int tid = threadIdx.x;
    const bool active = tid < ta * tc;
    const int a0 = active ? 4 * (tid / tc) : 0, c0 = active ? 4 * (tid % tc) : 0;
    float acc[4][4] = {};
    const int p0 = blockIdx.x * chunk, p1 = min(NN, p0 + chunk);
    for (int i = tid; i < 64 * 65; i += 256) {
        sa[i / 65][i % 65] = 0.f;
        sb[i / 65][i % 65] = Bm ? 0.f : 1.f;
    }
    __syncthreads();
    for (int b = 0; b < B; ++b)
        for (int pt = p0; pt < p1; pt += 64) {
            const int np_ = min(64, p1 - pt);
            for (int i = tid; i < na * 64; i += 256) {
                const int a = i >> 6, pp = i & 63;
                sa[a][pp] = pp < np_ ? A[((size_t)b * na + a) * NN + pt + pp] : 0.f;
            }
            if (Bm)
                for (int i = tid; i < nc * 64; i += 256) {
                    const int c = i >> 6, pp = i & 63;
                    sb[c][pp] = pp < np_ ? Bm[((size_t)b * nc + c) * NN + pt + pp] : 0.f;
                }
            __syncthreads();
            if (active) {
#pragma unroll 4
                for (int pp = 0; pp < 64; ++pp) {
                    float va[4], vb[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        va[r] = sa[min(a0 + r, 63)][pp];
                        vb[r] = sb[min(c0 + r, 63)][pp];
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int s = 0; s < 4; ++s) acc[r][s] += va[r] * vb[s];
                }
            }
            __syncthreads();
        }
    if (active)
        for (int r = 0; r < 4; ++r)
            for (int s = 0; s < 4; ++s)
                if (a0 + r < na && c0 + s < ncc)
                    part[((size_t)blockIdx.x * na + a0 + r) * ncc + c0 + s] = (double)acc[r][s];
}

// out[i] = sum over the chunks, in order, of part[chunk][i]
__global__ void k_sum_chunks(const double* __restrict__ part, int nch, int n, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[(size_t)c * n + i];
    out[i] = (float)s;
}

// dR[i][o][m] = s(m) sum_b Yb[b][o][m] conj(X[b][i][m])   (Yb = D du)
template <int C, int M>
__global__ void k_rgrad(const c32* __restrict__ Yb, const c32* __restrict__ X, float* __restrict__ dR, int B,
                        float inv_n2)
{
    constexpr int NM = 2 * M * M;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)C * C * NM) return;
    const int m = (int)(t % NM), o = (int)((t / NM) % C), i = (int)(t / ((size_t)NM * C));
    float gr = 0.f, gi = 0.f;
    for (int b = 0; b < B; ++b) {
        const c32 yb = Yb[((size_t)b * C + o) * NM + m], x = X[((size_t)b * C + i) * NM + m];
        gr += yb.x * x.x + yb.y * x.y;
        gi += yb.y * x.x - yb.x * x.y;
    }
    const float s = (m / (2 * M) == 0 ? 1.f : 2.f) * inv_n2;
    dR[2 * t] = s * gr;
    dR[2 * t + 1] = s * gi;
}

// Adam on the FP32 masters, refreshed BF16 copies
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wn = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wn;
    wb[i] = f2bf(wn);
}

// masters rounded to BF16, BF16 copies
__global__ void k_to_bf16(float* __restrict__ w, uint16_t* __restrict__ wb, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t h = f2bf(w[i]);
    wb[i] = h;
    w[i] = bf2f(h);
}

__global__ void k_fill0(float* __restrict__ p, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// planar FP32 (re, im) planes -> interleaved FP64 psi; real_only drops Im;
// base (may be null) is added to Re psi
__global__ void k_to_psi(const float* __restrict__ out, const float* __restrict__ base, double* __restrict__ psi,
                         size_t NN, int real_only)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= NN) return;
    psi[2 * i] = (double)out[i] + (base ? (double)base[i] : 0.0);
    psi[2 * i + 1] = real_only ? 0.0 : (double)out[NN + i];
}

// Periodic coordinate of the particle potential (src/gpe_particle.f90:31-45):
// X(dx) = (1/c)(s + s^3/6 + s^5/40), s = sin(c dx), c = 1/2 on the 2 pi box
__device__ inline double pcoord(double d)
{
    const double s = sin(0.5 * d), s2 = s * s;
    return 2.0 * (s + s * s2 * (1.0 / 6.0) + s * s2 * s2 * (1.0 / 40.0));
}

// Input planes of one sample: x[0] = sum_v q_v exp(-r_v^2 / (2 sigma^2)),
// x[1] = (V0 / vnorm) sum_p exp(-r_p^2 / (2 d^2)), r from pcoord per axis,
// and with cin >= 3 x[2] = sqrt(max(0, (omega - V)/g)) (Thomas-Fermi).
// src: xv[nv], yv[nv], qv[nv] (as doubles), pos[2*nobs]
__global__ void k_features(const double* __restrict__ src, int nv, int nobs, int N, double sigma, double amp,
                           double d, double vnorm, double omega, double g, int cin, float* __restrict__ x)
{
    const size_t NN = (size_t)N * N;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= NN) return;
    const double h = 2.0 * 3.14159265358979323846 / N;
    const double px = (double)(i % N) * h, py = (double)(i / N) * h;
    const double* xv = src;
    const double* yv = src + nv;
    const double* qv = src + 2 * nv;
    const double* pos = src + 3 * nv;
    const double is2 = 0.5 / (sigma * sigma), id2 = 0.5 / (d * d);
    double c = 0.0, v = 0.0;
    for (int k = 0; k < nv; ++k) {
        const double X = pcoord(px - xv[k]), Y = pcoord(py - yv[k]);
        c += qv[k] * exp(-(X * X + Y * Y) * is2);
    }
    for (int k = 0; k < nobs; ++k) {
        const double X = pcoord(px - pos[2 * k]), Y = pcoord(py - pos[2 * k + 1]);
        v += exp(-(X * X + Y * Y) * id2);
    }
    x[i] = (float)c;
    x[NN + i] = (float)(amp * v);
    if (cin >= 3) x[2 * NN + i] = (float)sqrt(fmax(0.0, (omega - vnorm * amp * v) / g));
    for (int k = 3; k < cin; ++k) x[k * NN + i] = 0.f;
}

// Thomas-Fermi guess from the potential plane: psi = sqrt(max(0, (omega - V)/g)), V = vnorm x[1]
__global__ void k_tf(const float* __restrict__ x, size_t NN, double omega, double g, double vnorm,
                     double* __restrict__ psi)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= NN) return;
    const double V = vnorm * (double)x[NN + i];
    psi[2 * i] = sqrt(fmax(0.0, (omega - V) / g));
    psi[2 * i + 1] = 0.0;
}

unsigned nblk(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

constexpr int FEAT_MAXV = 4096, FEAT_MAXO = 256;

}  // namespace

// ------------------------------------------------------------------ host --

struct vk_fno {
    int N = 0, cin = 0, C = 0, M = 0, L = 0, P = 0, B = 0;
    hipStream_t st = nullptr;
    size_t np = 0;
    size_t oWp = 0, obp = 0, oR[8] = {}, oW[8] = {}, ob[8] = {}, oQ1 = 0, obq1 = 0, oQ2 = 0, obq2 = 0;
    float *w = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;
    uint16_t* wb = nullptr;
    c32 *F = nullptr, *G = nullptr;
    float *x = nullptr, *tgt = nullptr, *out = nullptr, *dout = nullptr;
    float *z[9] = {}, *u[9] = {};
    c32 *A = nullptr, *X[8] = {}, *Y = nullptr, *Bk = nullptr, *Yb = nullptr;
    float *h = nullptr, *gq = nullptr, *dh = nullptr, *dz = nullptr, *du = nullptr;
    double* part = nullptr;
    double* psi = nullptr;
    double* fsrc = nullptr;   // staging for the feature builder
    long step = 0;
    size_t NN() const { return (size_t)N * N; }
    int NM() const { return 2 * M * M; }
};

namespace {

constexpr int OUT_CHUNK = 1024;   // pixels per weight-gradient block

int fno_alloc(vk_fno* f)
{
    const size_t NN = f->NN(), B = f->B, C = f->C, NM = f->NM();
    size_t o = 0;
    f->oWp = o, o += C * f->cin;
    f->obp = o, o += C;
    for (int l = 0; l < f->L; ++l) {
        f->oR[l] = o, o += 2 * C * C * NM;
        f->oW[l] = o, o += C * C;
        f->ob[l] = o, o += C;
    }
    f->oQ1 = o, o += (size_t)f->P * C;
    f->obq1 = o, o += f->P;
    f->oQ2 = o, o += 2 * (size_t)f->P;
    f->obq2 = o, o += 2;
    f->np = o;
    FCHK(hipMalloc(&f->w, o * 4));
    FCHK(hipMalloc(&f->g, o * 4));
    FCHK(hipMalloc(&f->am, o * 4));
    FCHK(hipMalloc(&f->av, o * 4));
    FCHK(hipMalloc(&f->wb, o * 2));
    FCHK(hipMemset(f->w, 0, o * 4));
    FCHK(hipMemset(f->wb, 0, o * 2));
    FCHK(hipMemset(f->g, 0, o * 4));
    FCHK(hipMemset(f->am, 0, o * 4));
    FCHK(hipMemset(f->av, 0, o * 4));
    FCHK(hipMalloc(&f->x, B * f->cin * NN * 4));
    FCHK(hipMalloc(&f->tgt, B * 2 * NN * 4));
    FCHK(hipMalloc(&f->out, B * 2 * NN * 4));
    FCHK(hipMalloc(&f->dout, B * 2 * NN * 4));
    for (int l = 0; l <= f->L; ++l) {
        FCHK(hipMalloc(&f->z[l], B * C * NN * 4));
        FCHK(hipMalloc(&f->u[l], B * C * NN * 4));
    }
    FCHK(hipMalloc(&f->A, B * C * f->N * f->M * 8));
    for (int l = 0; l < f->L; ++l) FCHK(hipMalloc(&f->X[l], B * C * NM * 8));
    FCHK(hipMalloc(&f->Y, B * C * NM * 8));
    FCHK(hipMalloc(&f->Yb, B * C * NM * 8));
    FCHK(hipMalloc(&f->Bk, B * C * f->N * f->M * 8));
    FCHK(hipMalloc(&f->h, B * f->P * NN * 4));
    FCHK(hipMalloc(&f->gq, B * f->P * NN * 4));
    FCHK(hipMalloc(&f->dh, B * f->P * NN * 4));
    FCHK(hipMalloc(&f->dz, B * C * NN * 4));
    FCHK(hipMalloc(&f->du, B * C * NN * 4));
    FCHK(hipMalloc(&f->psi, NN * 16));
    FCHK(hipMalloc(&f->fsrc, (3 * FEAT_MAXV + 2 * FEAT_MAXO) * sizeof(double)));
    const size_t nch = (NN + OUT_CHUNK - 1) / OUT_CHUNK;
    FCHK(hipMalloc(&f->part, std::max<size_t>(nch * 64 * 64, 1024) * 8));
    // e^{-2 pi i k x / N}; the phase index k*x is reduced mod N exactly
    std::vector<c32> F((size_t)f->N * f->M), G((size_t)f->N * 2 * f->M);
    const double pi = 3.14159265358979323846;
    for (int x = 0; x < f->N; ++x) {
        for (int k = 0; k < f->M; ++k) {
            const double a = -2.0 * pi * (double)(((long)k * x) % f->N) / f->N;
            F[(size_t)x * f->M + k] = make_float2((float)cos(a), (float)sin(a));
        }
        for (int j = 0; j < 2 * f->M; ++j) {
            const long ky = j < f->M ? j : j - 2 * f->M;
            const long kk = ((ky * x) % f->N + f->N) % f->N;
            const double a = -2.0 * pi * (double)kk / f->N;
            G[(size_t)x * 2 * f->M + j] = make_float2((float)cos(a), (float)sin(a));
        }
    }
    FCHK(hipMalloc(&f->F, F.size() * 8));
    FCHK(hipMalloc(&f->G, G.size() * 8));
    FCHK(hipMemcpy(f->F, F.data(), F.size() * 8, hipMemcpyHostToDevice));
    FCHK(hipMemcpy(f->G, G.data(), G.size() * 8, hipMemcpyHostToDevice));
    return 0;
}

template <int C, int M>
struct Net {
    static void dft(vk_fno* f, const float* z, c32* X, int nq)
    {
        hipLaunchKernelGGL(k_dftx<M>, dim3(f->N / 16, nq), dim3(16 * M), 0, f->st, z, (const c32*)f->F, f->A, f->N);
        hipLaunchKernelGGL(k_dfty<M>, dim3(nq), dim3(2 * M * M), 0, f->st, (const c32*)f->A, (const c32*)f->G, X,
                           f->N);
    }

    static void spectral_inv(vk_fno* f, int nb)
    {
        hipLaunchKernelGGL(k_idfty<M>, dim3(nblk((size_t)nb * C * f->N * M, 256)), dim3(256), 0, f->st,
                           (const c32*)f->Y, (const c32*)f->G, f->Bk, f->N, nb * C);
    }

    static void outer(vk_fno* f, const float* A, const float* Bm, int nb, int na, int nc, size_t off)
    {
        const int NN = (int)f->NN();
        const int nch = (NN + OUT_CHUNK - 1) / OUT_CHUNK;
        const int ncc = Bm ? nc : 1;
        hipLaunchKernelGGL(k_outer, dim3(nch), dim3(256), 0, f->st, A, Bm, f->part, nb, na, nc, NN, OUT_CHUNK);
        hipLaunchKernelGGL(k_sum_chunks, dim3(nblk((size_t)na * ncc, 256)), dim3(256), 0, f->st,
                           (const double*)f->part, nch, na * ncc, f->g + off);
    }

    static void forward(vk_fno* f, int nb)
    {
        const int NN = (int)f->NN();
        const float inv_n2 = 1.0f / ((float)f->N * (float)f->N);
        hipLaunchKernelGGL(k_lift<C>, dim3(nblk((size_t)nb * NN, 256)), dim3(256), 0, f->st, (const float*)f->x,
                           f->wb + f->oWp, f->wb + f->obp, f->z[0], nb, f->cin, NN);
        for (int l = 0; l < f->L; ++l) {
            dft(f, f->z[l], f->X[l], nb * C);
            hipLaunchKernelGGL((k_mix<C, M>), dim3(nblk((size_t)nb * C * 2 * M * M, 256)), dim3(256), 0, f->st,
                               (const c32*)f->X[l], f->wb + f->oR[l], f->Y, nb, 0, inv_n2);
            spectral_inv(f, nb);
            hipLaunchKernelGGL((k_out<C, M>), dim3(f->N, nb), dim3(256), 0, f->st, (const float*)f->z[l],
                               f->wb + f->oW[l], f->wb + f->ob[l], (const c32*)f->Bk, (const c32*)f->F, f->u[l + 1],
                               f->z[l + 1], (const float*)nullptr, f->N, 0, l < f->L - 1 ? 1 : 0);
        }
        hipLaunchKernelGGL(k_proj<C>, dim3(nblk((size_t)nb * NN, 256)), dim3(256), 0, f->st, (const float*)f->z[f->L],
                           f->wb + f->oQ1, f->wb + f->obq1, f->wb + f->oQ2, f->wb + f->obq2, f->h, f->out, nb, f->P,
                           NN);
    }

    static void backward(vk_fno* f, int nb)
    {
        const int NN = (int)f->NN(), P = f->P;
        const float inv_n2 = 1.0f / ((float)f->N * (float)f->N);
        hipLaunchKernelGGL(k_proj_bwd<C>, dim3(nblk((size_t)nb * NN, 256)), dim3(256), 0, f->st, (const float*)f->h,
                           (const float*)f->dout, f->wb + f->oQ1, f->wb + f->oQ2, f->dh, f->gq, f->dz, nb, P, NN);
        outer(f, f->dout, f->gq, nb, 2, P, f->oQ2);
        outer(f, f->dout, nullptr, nb, 2, 1, f->obq2);
        outer(f, f->dh, f->z[f->L], nb, P, C, f->oQ1);
        outer(f, f->dh, nullptr, nb, P, 1, f->obq1);
        float* dcur = f->dz;   // dL/du_{l+1}; the last layer has no activation
        float* dnext = f->du;
        for (int l = f->L - 1; l >= 0; --l) {
            outer(f, dcur, f->z[l], nb, C, C, f->oW[l]);
            outer(f, dcur, nullptr, nb, C, 1, f->ob[l]);
            dft(f, dcur, f->Yb, nb * C);
            hipLaunchKernelGGL((k_rgrad<C, M>), dim3(nblk((size_t)C * C * 2 * M * M, 256)), dim3(256), 0, f->st,
                               (const c32*)f->Yb, (const c32*)f->X[l], f->g + f->oR[l], nb, inv_n2);
            hipLaunchKernelGGL((k_mix<C, M>), dim3(nblk((size_t)nb * C * 2 * M * M, 256)), dim3(256), 0, f->st,
                               (const c32*)f->Yb, f->wb + f->oR[l], f->Y, nb, 1, inv_n2);
            spectral_inv(f, nb);
            hipLaunchKernelGGL((k_out<C, M>), dim3(f->N, nb), dim3(256), 0, f->st, (const float*)dcur,
                               f->wb + f->oW[l], (const uint16_t*)nullptr, (const c32*)f->Bk, (const c32*)f->F,
                               (float*)nullptr, dnext, l > 0 ? (const float*)f->u[l] : (const float*)nullptr, f->N,
                               1, 0);
            std::swap(dcur, dnext);
        }
        outer(f, dcur, f->x, nb, C, f->cin, f->oWp);
        outer(f, dcur, nullptr, nb, C, 1, f->obp);
    }
};

typedef void (*pass_fn)(vk_fno*, int);

int dispatch(const vk_fno* f, pass_fn* fw, pass_fn* bw)
{
    if (f->C == 32 && f->M == 16) {
        *fw = Net<32, 16>::forward, *bw = Net<32, 16>::backward;
    } else if (f->C == 16 && f->M == 8) {
        *fw = Net<16, 8>::forward, *bw = Net<16, 8>::backward;
    } else if (f->C == 8 && f->M == 4) {
        *fw = Net<8, 4>::forward, *bw = Net<8, 4>::backward;
    } else {
        return vk_set_error(VK_ERR_UNSUPPORTED, "fno: (width, modes) must be (32,16), (16,8) or (8,4)");
    }
    return 0;
}

int load_input(vk_fno* f, const float* x, int nb)
{
    FCHK(hipMemcpyAsync(f->x, x, (size_t)nb * f->cin * f->NN() * 4, hipMemcpyDefault, f->st));
    return 0;
}

}  // namespace

extern "C" {

int vk_fno_create(vk_fno** out, int N, int cin, int width, int modes, int layers, int proj, int batch)
{
    if (!out) return vk_set_error(VK_ERR_ARG, "fno_create: null handle pointer");
    *out = nullptr;
    if (N < 64 || N > 1024 || N % 64 || cin < 1 || cin > CIN_MAX || layers < 1 || layers > 8 || proj < 1 ||
        proj > 64 || batch < 1 || batch > 64 || 2 * modes > N)
        return vk_set_error(VK_ERR_ARG, "fno_create: N must be a multiple of 64 in [64, 1024]; cin <= 4, "
                                        "layers <= 8, proj <= 64, batch <= 64, 2*modes <= N");
    vk_fno probe;
    probe.C = width, probe.M = modes;
    pass_fn a, b;
    int rc = dispatch(&probe, &a, &b);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return vk_set_error(VK_ERR_HIP, "no ROCm device available");
    vk_fno* f = new vk_fno();
    f->N = N, f->cin = cin, f->C = width, f->M = modes, f->L = layers, f->P = proj, f->B = batch;
    if (hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking) != hipSuccess) {
        delete f;
        return vk_set_error(VK_ERR_HIP, "fno: stream creation failed");
    }
    if ((rc = fno_alloc(f))) {
        vk_fno_destroy(f);
        return rc;
    }
    *out = f;
    return 0;
}

void vk_fno_destroy(vk_fno* f)
{
    if (!f) return;
    if (f->st) hipStreamSynchronize(f->st);
    void* ps[] = {f->w, f->g, f->am, f->av, f->wb, f->F, f->G, f->x, f->tgt, f->out, f->dout, f->A, f->Y, f->Yb,
                  f->Bk, f->h, f->gq, f->dh, f->dz, f->du, f->part, f->psi, f->fsrc};
    for (void* p : ps)
        if (p) hipFree(p);
    for (float* p : f->z)
        if (p) hipFree(p);
    for (float* p : f->u)
        if (p) hipFree(p);
    for (c32* p : f->X)
        if (p) hipFree(p);
    if (f->st) hipStreamDestroy(f->st);
    delete f;
}

long vk_fno_nparams(const vk_fno* f) { return f ? (long)f->np : 0; }

int vk_fno_layout(const vk_fno* f, long* offsets)
{
    if (!f || !offsets) return vk_set_error(VK_ERR_ARG, "fno_layout: null");
    int k = 0;
    offsets[k++] = (long)f->oWp;
    offsets[k++] = (long)f->obp;
    for (int l = 0; l < f->L; ++l) {
        offsets[k++] = (long)f->oR[l];
        offsets[k++] = (long)f->oW[l];
        offsets[k++] = (long)f->ob[l];
    }
    offsets[k++] = (long)f->oQ1;
    offsets[k++] = (long)f->obq1;
    offsets[k++] = (long)f->oQ2;
    offsets[k++] = (long)f->obq2;
    offsets[k++] = (long)f->np;
    return k;
}

void* vk_fno_stream(const vk_fno* f) { return f ? (void*)f->st : nullptr; }

int vk_fno_set_params(vk_fno* f, const float* p)
{
    if (!f || !p) return vk_set_error(VK_ERR_ARG, "fno_set_params: null");
    FCHK(hipMemcpyAsync(f->w, p, f->np * 4, hipMemcpyDefault, f->st));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(f->np, 256)), dim3(256), 0, f->st, f->w, f->wb, f->np);
    hipLaunchKernelGGL(k_fill0, dim3(nblk(f->np, 256)), dim3(256), 0, f->st, f->am, f->np);
    hipLaunchKernelGGL(k_fill0, dim3(nblk(f->np, 256)), dim3(256), 0, f->st, f->av, f->np);
    f->step = 0;
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipGetLastError());
    return 0;
}

int vk_fno_get_params(vk_fno* f, float* p)
{
    if (!f || !p) return vk_set_error(VK_ERR_ARG, "fno_get_params: null");
    std::vector<uint16_t> b(f->np);
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipMemcpy(b.data(), f->wb, f->np * 2, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < f->np; ++i) p[i] = bf2f(b[i]);
    return 0;
}

int vk_fno_get_grad(vk_fno* f, float* gout)
{
    if (!f || !gout) return vk_set_error(VK_ERR_ARG, "fno_get_grad: null");
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipMemcpy(gout, f->g, f->np * 4, hipMemcpyDefault));
    return 0;
}

int vk_fno_forward(vk_fno* f, const float* x, float* y, int nb)
{
    if (!f || !x || nb < 1 || nb > f->B) return vk_set_error(VK_ERR_ARG, "fno_forward: bad arguments");
    pass_fn fw, bw;
    int rc = dispatch(f, &fw, &bw);
    if (rc || (rc = load_input(f, x, nb))) return rc;
    fw(f, nb);
    if (y) FCHK(hipMemcpyAsync(y, f->out, (size_t)nb * 2 * f->NN() * 4, hipMemcpyDefault, f->st));
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipGetLastError());
    return 0;
}

int vk_fno_predict_psi(vk_fno* f, const float* x, double* psi, int flags)
{
    if (!f || !psi || (flags & ~(VK_FNO_REAL | VK_FNO_ADD_TF)) || ((flags & VK_FNO_ADD_TF) && f->cin < 3))
        return vk_set_error(VK_ERR_ARG, "fno_predict_psi: bad arguments");
    pass_fn fw, bw;
    int rc = dispatch(f, &fw, &bw);
    if (rc || (x && (rc = load_input(f, x, 1)))) return rc;
    fw(f, 1);
    const float* base = (flags & VK_FNO_ADD_TF) ? f->x + 2 * f->NN() : nullptr;
    hipLaunchKernelGGL(k_to_psi, dim3(nblk(f->NN(), 256)), dim3(256), 0, f->st, (const float*)f->out, base, f->psi,
                       f->NN(), flags & VK_FNO_REAL);
    FCHK(hipMemcpyAsync(psi, f->psi, f->NN() * 16, hipMemcpyDefault, f->st));
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipGetLastError());
    return 0;
}

int vk_fno_features(vk_fno* f, int nv, const double* xv, const double* yv, const int* qv, double sigma, int nobs,
                    const double* pos, double V0, double d, double vnorm, double omega, double g, float* x_out)
{
    if (!f || f->cin < 2 || nv < 0 || nv > FEAT_MAXV || nobs < 0 || nobs > FEAT_MAXO || (nv && (!xv || !yv || !qv)) ||
        (nobs && (!pos || !(d > 0.0))) || !(sigma > 0.0) || !(vnorm > 0.0) || !(g > 0.0))
        return vk_set_error(VK_ERR_ARG, "fno_features: bad arguments (nv <= 4096, nobs <= 256, cin >= 2, "
                                        "sigma, vnorm, g and d > 0)");
    std::vector<double> src(3 * (size_t)nv + 2 * (size_t)nobs);
    for (int k = 0; k < nv; ++k) {
        src[k] = xv[k];
        src[nv + k] = yv[k];
        src[2 * nv + k] = (double)qv[k];
    }
    for (int k = 0; k < 2 * nobs; ++k) src[3 * nv + k] = pos[k];
    FCHK(hipStreamSynchronize(f->st));
    if (!src.empty()) FCHK(hipMemcpy(f->fsrc, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_features, dim3(nblk(f->NN(), 256)), dim3(256), 0, f->st, (const double*)f->fsrc, nv, nobs,
                       f->N, sigma, V0 / vnorm, nobs ? d : 1.0, vnorm, omega, g, f->cin, f->x);
    if (x_out) FCHK(hipMemcpyAsync(x_out, f->x, (size_t)f->cin * f->NN() * 4, hipMemcpyDefault, f->st));
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipGetLastError());
    return 0;
}

int vk_fno_tf_guess(vk_fno* f, double omega, double g, double vnorm, double* psi)
{
    if (!f || !psi || f->cin < 2 || !(g > 0.0)) return vk_set_error(VK_ERR_ARG, "fno_tf_guess: bad arguments");
    hipLaunchKernelGGL(k_tf, dim3(nblk(f->NN(), 256)), dim3(256), 0, f->st, (const float*)f->x, f->NN(), omega, g,
                       vnorm, f->psi);
    FCHK(hipMemcpyAsync(psi, f->psi, f->NN() * 16, hipMemcpyDefault, f->st));
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipGetLastError());
    return 0;
}

int vk_fno_train_step(vk_fno* f, const float* x, const float* tgt, int nb, float lr, double* loss)
{
    if (!f || !x || !tgt || nb < 1 || nb > f->B || !(lr >= 0.f))
        return vk_set_error(VK_ERR_ARG, "fno_train_step: bad arguments");
    pass_fn fw, bw;
    int rc = dispatch(f, &fw, &bw);
    if (rc || (rc = load_input(f, x, nb))) return rc;
    const size_t n = (size_t)nb * 2 * f->NN();
    FCHK(hipMemcpyAsync(f->tgt, tgt, n * 4, hipMemcpyDefault, f->st));
    fw(f, nb);
    constexpr int NBL = 1024;
    hipLaunchKernelGGL(k_loss, dim3(NBL), dim3(256), 0, f->st, (const float*)f->out, (const float*)f->tgt, f->dout,
                       f->part, n, 1.0f / (float)n);
    std::vector<double> lp(NBL);
    FCHK(hipMemcpyAsync(lp.data(), f->part, NBL * 8, hipMemcpyDeviceToHost, f->st));
    FCHK(hipStreamSynchronize(f->st));
    double L = 0.0;
    for (double v : lp) L += v;
    if (loss) *loss = L;
    bw(f, nb);
    if (lr > 0.f) {
        f->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)f->step), c2 = 1.f - powf(b2, (float)f->step);
        hipLaunchKernelGGL(k_adam, dim3(nblk(f->np, 256)), dim3(256), 0, f->st, f->w, f->wb, (const float*)f->g,
                           f->am, f->av, f->np, lr, b1, b2, 1e-8f, c1, c2);
    }
    FCHK(hipStreamSynchronize(f->st));
    FCHK(hipGetLastError());
    return 0;
}

}  // extern "C"
