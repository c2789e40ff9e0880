// Batched ensemble of independent realizations of unequal size.  One launch
// per pass serves every realization: each workgroup finds its realization
// through a block map and runs the size's pass body.  Sizes 128..1024 all
// use 256-thread workgroups, so one kernel holds every size.  P1 stores the
// x-space psi rows it computes (Xs) and P3 reads them, which saves P3 one of
// its three transforms; the rows are the bits P3 would compute itself.
#pragma once
#include "vk_kernels.hpp"

// Occupancy of the ensemble passes (waves per SIMD; <= 168 VGPRs for 3):
// three 256-thread workgroups per CU for the row passes (P1, P3) and the
// column passes (P2, P4) make the 384-member step 4.7% faster than the
// compiler's own register budget (two to three), and four are slower.
#ifndef VK_ENS_ROW_WAVES
#define VK_ENS_ROW_WAVES 3
#endif
#ifndef VK_ENS_COL_WAVES
#define VK_ENS_COL_WAVES 3
#endif
#define VK_ENS_ROCC __attribute__((amdgpu_waves_per_eu(VK_ENS_ROW_WAVES)))
#define VK_ENS_COCC __attribute__((amdgpu_waves_per_eu(VK_ENS_COL_WAVES)))

namespace vk {

struct EnsReal {
    int N;
    int var;        // shape variant (CfgE<N, var>)
    Geo g;
    cplx *K0, *K1, *K2, *Hpsi, *Hrho, *HN, *Xs;
    const cplx* tw;
    double alpha, omega, gsc, sc, dt;
    cplx pref;
};

#define VK_ENS_SWITCH(N, BODY)                                           \
    switch (N * 2 + e.var) {                                              \
    case 256: { using C = CfgE<128, 0>; BODY; break; }                    \
    case 512: { using C = CfgE<256, 0>; BODY; break; }                    \
    case 513: { using C = CfgE<256, 1>; BODY; break; }                    \
    case 1024: { using C = CfgE<512, 0>; BODY; break; }                   \
    case 1025: { using C = CfgE<512, 1>; BODY; break; }                   \
    case 2048: { using C = CfgE<1024, 0>; BODY; break; }                  \
    case 2049: { using C = CfgE<1024, 1>; BODY; break; }                  \
    default: break;                                                       \
    }

#define VK_ENS_SETUP                                                      \
    const int r = map[blockIdx.x];                                        \
    const EnsReal& e = R[r];                                              \
    const int lb = blockIdx.x - off[2 * r], nb = off[2 * r + 1] - off[2 * r];   // [start, end) of r

__global__ void __launch_bounds__(256) VK_ENS_ROCC k_ens_p1(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off)
{
    VK_ENS_SETUP
    VK_ENS_SWITCH(e.N, (p1_body<C, true>(lb, nb, e.Hpsi, e.Hrho, e.tw, e.g, 1, e.sc, e.Xs)))
}

__global__ void __launch_bounds__(256) VK_ENS_COCC k_ens_p2(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off)
{
    VK_ENS_SETUP
    VK_ENS_SWITCH(e.N, (p2_body<C>(lb, nb, e.Hrho, e.Hrho, (const cplx*)e.K1, e.K1, e.tw, e.g, 0)))
}

__global__ void __launch_bounds__(256) VK_ENS_ROCC k_ens_p3(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off)
{
    VK_ENS_SETUP
    const P3Part pa{};
    VK_ENS_SWITCH(e.N, (p3_body<C, 0, true>(lb, nb, e.Hrho, e.Hrho, nullptr, e.HN, e.tw, e.g, 0, 1, e.gsc, e.sc,
                                            1.0, 1.0, pa, e.Xs)))
}

__global__ void __launch_bounds__(256) VK_ENS_COCC k_ens_p4(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off, int stage)
{
    VK_ENS_SETUP
    RKArgs a{};
    a.pref = e.pref;
    a.alpha = e.alpha;
    a.omega = e.omega;
    a.stage = stage;
    const double oney6 = 1.0 / 6.0;
    a.wacc = (stage == 1 || stage == 4) ? e.dt * oney6 : e.dt * oney6 * 2.0;
    a.cstep = stage == 3 ? e.dt : e.dt * 0.5;
    a.r1 = (e.dt * oney6) / (e.dt * 0.5);
    a.r3 = (e.dt * oney6 * 2.0) / e.dt;
    switch (stage) {
    case 1: VK_ENS_SWITCH(e.N, (p4_body<C, 1>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    case 2: VK_ENS_SWITCH(e.N, (p4_body<C, 2>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    case 3: VK_ENS_SWITCH(e.N, (p4_body<C, 3>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    default: VK_ENS_SWITCH(e.N, (p4_body<C, 4>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    }
}

}  // namespace vk
