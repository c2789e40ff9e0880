// Attainable HBM bandwidth of the passes' access patterns, measured on the
// device: the ceilings the pass kernels' roofline fractions are read
// against.  Each pattern moves one 4096^2 complex128 field (256 MiB) in and
// one out, with no arithmetic, over the passes' panel layout
// H[(p*N + y)*TC + c] with TC = Cfg<4096>::TC:
//   0  contiguous copy;
//   1  row-pass pattern: a 256-thread workgroup per row gathers the row's
//      pieces (TC points) of every panel and writes the row contiguously
//      (P1 and P3 with one row per workgroup);
//   2  column-pass pattern: a workgroup of 256 threads per column copies
//      the P4 share of a panel (the P4 columns per workgroup), sibling
//      workgroups of a panel on one XCD;
//   3  the same with one column per 256-thread workgroup (P2).
#include "vk_cfg.hpp"
#include "vikshobha.h"

#include <string>

int vk_set_error(int code, const std::string& msg);

namespace {

using vk::cplx;
constexpr int PN = 4096, PTC = vk::Cfg<4096>::TC;
#ifndef VK_P4_COLS_4096
#define VK_P4_COLS_4096 2
#endif
constexpr int PG4 = VK_P4_COLS_4096 > 0 ? VK_P4_COLS_4096 : PTC;

__global__ void k_pat_copy(const cplx* __restrict__ a, cplx* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void __launch_bounds__(256) k_pat_row(const cplx* __restrict__ H, cplx* __restrict__ X)
{
    constexpr int T = 256, S = 16, TR = 1;
    const int u = threadIdx.x;
    const int q = u % (TR * PTC), rr = q / PTC, c = q % PTC, pidx0 = u / (TR * PTC);
    const int y = vk::xcd_remap_blk(blockIdx.x, gridDim.x) * TR + rr;
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = H[((size_t)(pidx0 + (T / PTC) * s) * PN + y) * PTC + c];
#pragma unroll
    for (int s = 0; s < S; ++s) X[(size_t)y * PN + PTC * pidx0 + c + T * s] = v[s];
}

// G columns of each TC-column panel per workgroup
template <int G>
__global__ void __launch_bounds__(G * 256) k_pat_col(const cplx* __restrict__ H, cplx* __restrict__ K)
{
    constexpr int T = 256, S = 16, GS = PTC / G;
    int pb = blockIdx.x, hh = 0;
    if (GS > 1) {
        const int r = pb % (8 * GS);
        hh = r / 8;
        pb = pb / (8 * GS) * 8 + r % 8;
    }
    const int u = threadIdx.x, c = hh * G + u % G, t = u / G;
    cplx v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) v[s] = H[((size_t)pb * PN + t + T * s) * PTC + c];
#pragma unroll
    for (int s = 0; s < S; ++s) K[((size_t)pb * PN + t + T * s) * PTC + c] = v[s];
}

}  // namespace

// GB/s (bytes read + written per launch / average launch time over reps
// launches) of patterns 0..3 into gbps[0..n-1]; two 256 MiB buffers are
// allocated for the measurement and freed.
extern "C" int vk_pattern_bandwidth(int reps, double* gbps, int n)
{
    if (reps <= 0 || !gbps || n <= 0) return vk_set_error(VK_ERR_ARG, "pattern_bandwidth: bad arguments");
    const size_t cnt = (size_t)PN * PN;
    cplx *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, cnt * sizeof(cplx)) != hipSuccess || hipMalloc(&b, cnt * sizeof(cplx)) != hipSuccess) {
        if (a) (void)hipFree(a);
        return vk_set_error(VK_ERR_HIP, "pattern_bandwidth: hipMalloc");
    }
    (void)hipMemset(a, 0, cnt * sizeof(cplx));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    constexpr int NP = PN / PTC;
    for (int k = 0; k < n && k < 4; ++k) {
        auto launch = [&]() {
            switch (k) {
            case 0: hipLaunchKernelGGL(k_pat_copy, dim3(8192), dim3(256), 0, 0, a, b, cnt); break;
            case 1: hipLaunchKernelGGL(k_pat_row, dim3(PN), dim3(256), 0, 0, a, b); break;
            case 2: hipLaunchKernelGGL(k_pat_col<PG4>, dim3(NP * (PTC / PG4)), dim3(PG4 * 256), 0, 0, a, b); break;
            default: hipLaunchKernelGGL(k_pat_col<1>, dim3(NP * PTC), dim3(256), 0, 0, a, b); break;
            }
        };
        for (int i = 0; i < 2; ++i) launch();
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < reps; ++i) launch();
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        gbps[k] = 2.0 * cnt * sizeof(cplx) / ((double)ms / reps * 1e-3) / 1e9;
    }
    const hipError_t err = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return vk_set_error(VK_ERR_HIP, std::string("pattern_bandwidth: ") + hipGetErrorString(err));
    return 0;
}
