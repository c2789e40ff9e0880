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

#define VK_ENS_KEYS(KEY, BODY)                                           \
    switch (KEY) {                                                        \
    case 256: { using C = CfgE<128, 0>; BODY; break; }                    \
    case 512: { using C = CfgE<256, 0>; BODY; break; }                    \
    case 513: { using C = CfgE<256, 1>; BODY; break; }                    \
    case 1024: { using C = CfgE<512, 0>; BODY; break; }                   \
    case 1025: { using C = CfgE<512, 1>; BODY; break; }                   \
    case 2048: { using C = CfgE<1024, 0>; BODY; break; }                  \
    case 2049: { using C = CfgE<1024, 1>; BODY; break; }                  \
    default: break;                                                       \
    }

#define VK_ENS_SWITCH(N, BODY) VK_ENS_KEYS((N) * 2 + e.var, BODY)

// the same over the loop shapes CfgL
#define VK_ENS_LKEYS(KEY, BODY)                                          \
    switch (KEY) {                                                        \
    case 256: { using C = CfgL<128, 0>; BODY; break; }                    \
    case 512: { using C = CfgL<256, 0>; BODY; break; }                    \
    case 513: { using C = CfgL<256, 1>; BODY; break; }                    \
    case 1024: { using C = CfgL<512, 0>; BODY; break; }                   \
    case 1025: { using C = CfgL<512, 1>; BODY; break; }                   \
    case 2048: { using C = CfgL<1024, 0>; BODY; break; }                  \
    case 2049: { using C = CfgL<1024, 1>; BODY; break; }                  \
    default: break;                                                       \
    }

// the member of block b of a pass's block map, and b's index in that member
#define VK_ENS_SETUP_B(b)                                                 \
    const int r = map[b];                                                 \
    const EnsReal& e = R[r];                                              \
    const int lb = (b) - off[2 * r], nb = off[2 * r + 1] - off[2 * r];   // [start, end) of r

// RK4 stage coefficients of member e
__device__ __forceinline__ RKArgs ens_rk_args(const EnsReal& e, int stage)
{
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
    return a;
}

// One block of each pass for the member that owns block b (a workgroup of
// 256 threads; the pass bodies hold their own workgroup barriers).
__device__ __forceinline__ void ens_tile_p1(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                            const int* __restrict__ off, int b)
{
    VK_ENS_SETUP_B(b)
    VK_ENS_SWITCH(e.N, (p1_body<C, true>(lb, nb, e.Hpsi, e.Hrho, e.tw, e.g, 1, e.sc, e.Xs)))
}

__device__ __forceinline__ void ens_tile_p2(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                            const int* __restrict__ off, int b)
{
    VK_ENS_SETUP_B(b)
    VK_ENS_SWITCH(e.N, (p2_body<C>(lb, nb, e.Hrho, e.Hrho, (const cplx*)e.K1, e.K1, e.tw, e.g, 0)))
}

__device__ __forceinline__ void ens_tile_p3(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                            const int* __restrict__ off, int b)
{
    VK_ENS_SETUP_B(b)
    const P3Part pa{};
    VK_ENS_SWITCH(e.N, (p3_body<C, 0, true>(lb, nb, e.Hrho, e.Hrho, nullptr, e.HN, e.tw, e.g, 0, 1, e.gsc, e.sc,
                                            1.0, 1.0, pa, e.Xs)))
}

__device__ __forceinline__ void ens_tile_p4(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                            const int* __restrict__ off, int b, int stage)
{
    VK_ENS_SETUP_B(b)
    const RKArgs a = ens_rk_args(e, stage);
    switch (stage) {
    case 1: VK_ENS_SWITCH(e.N, (p4_body<C, 1>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    case 2: VK_ENS_SWITCH(e.N, (p4_body<C, 2>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    case 3: VK_ENS_SWITCH(e.N, (p4_body<C, 3>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    default: VK_ENS_SWITCH(e.N, (p4_body<C, 4>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, a))) break;
    }
}

__global__ void __launch_bounds__(256) VK_ENS_ROCC k_ens_p1(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off)
{
    ens_tile_p1(R, map, off, blockIdx.x);
}

__global__ void __launch_bounds__(256) VK_ENS_COCC k_ens_p2(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off)
{
    ens_tile_p2(R, map, off, blockIdx.x);
}

__global__ void __launch_bounds__(256) VK_ENS_ROCC k_ens_p3(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off)
{
    ens_tile_p3(R, map, off, blockIdx.x);
}

__global__ void __launch_bounds__(256) VK_ENS_COCC k_ens_p4(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                const int* __restrict__ off, int stage)
{
    ens_tile_p4(R, map, off, blockIdx.x, stage);
}

// the same with the stage a compile-time constant: one stage's bodies per
// kernel (fewer registers than the four stages behind a switch)
template <int STAGE>
__global__ void __launch_bounds__(256) VK_ENS_COCC k_ens_p4s(const EnsReal* __restrict__ R, const int* __restrict__ map,
                                                   const int* __restrict__ off)
{
    VK_ENS_SETUP_B(blockIdx.x)
    VK_ENS_SWITCH(e.N, (p4_body<C, STAGE>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, ens_rk_args(e, STAGE))))
}

// ---------------------------------------------------------------------------
// Persistent step: one launch runs the 16 passes of an RK4 step for every
// member, with a grid barrier between passes.  The blocks of a pass are
// numbered over the member sizes in turn (one group per size, each with its
// own block map); workgroup w takes the blocks whose number is w modulo the
// grid.  The grid is at most the number of workgroups the device holds at
// once (all resident) and a multiple of 8, so a block runs on the XCD it
// would in a per-pass launch of the same numbering.  Each group's loop is
// its own code (the shape is a template argument): a loop over blocks of
// several shapes would hold every shape's loop-invariant values at once and
// spill.  A barrier waits at most `limit` wall-clock ticks: a workgroup that
// waits longer sets *err, every workgroup then leaves its waits, and the
// kernel ends (the step's results are invalid; the host reports it).

constexpr int kEnsMaxGroups = 4;       // member sizes 128, 256, 512, 1024
constexpr int kEnsStepBarriers = 15;   // between the 16 passes of a step

struct EnsGroup {
    int key;              // N * 2 + shape variant
    const int* map[3];    // the size's block maps: row passes, P2, P4
    const int* off[3];
    int nblk[3];
    int base[3];          // number of the group's first block in each pass
};

struct EnsStepArgs {
    EnsGroup grp[kEnsMaxGroups];
    int ngrp;
    unsigned* bar;        // one counter per barrier, zero at launch
    int* err;
    long long limit;
};

// this workgroup's first block of a group whose blocks start at number base
__device__ __forceinline__ int ens_first(int base)
{
    const int G = gridDim.x;
    return ((int)blockIdx.x - base % G + G) % G;
}

template <class C>
__device__ __forceinline__ void ens_loop_p1(const EnsReal* __restrict__ R, const EnsGroup& G)
{
    const int* __restrict__ map = G.map[0];
    const int* __restrict__ off = G.off[0];
#pragma unroll 1
    for (int b = ens_first(G.base[0]); b < G.nblk[0]; b += gridDim.x) {
        VK_ENS_SETUP_B(b)
        p1_body<C, true>(lb, nb, e.Hpsi, e.Hrho, e.tw, e.g, 1, e.sc, e.Xs);
        __syncthreads();
    }
}

template <class C>
__device__ __forceinline__ void ens_loop_p2(const EnsReal* __restrict__ R, const EnsGroup& G)
{
    const int* __restrict__ map = G.map[1];
    const int* __restrict__ off = G.off[1];
#pragma unroll 1
    for (int b = ens_first(G.base[1]); b < G.nblk[1]; b += gridDim.x) {
        VK_ENS_SETUP_B(b)
        p2_body<C>(lb, nb, e.Hrho, e.Hrho, (const cplx*)e.K1, e.K1, e.tw, e.g, 0);
        __syncthreads();
    }
}

template <class C>
__device__ __forceinline__ void ens_loop_p3(const EnsReal* __restrict__ R, const EnsGroup& G)
{
    const int* __restrict__ map = G.map[0];
    const int* __restrict__ off = G.off[0];
    const P3Part pa{};
#pragma unroll 1
    for (int b = ens_first(G.base[0]); b < G.nblk[0]; b += gridDim.x) {
        VK_ENS_SETUP_B(b)
        p3_body<C, 0, true>(lb, nb, e.Hrho, e.Hrho, nullptr, e.HN, e.tw, e.g, 0, 1, e.gsc, e.sc, 1.0, 1.0, pa, e.Xs);
        __syncthreads();
    }
}

template <class C, int STAGE>
__device__ __forceinline__ void ens_loop_p4(const EnsReal* __restrict__ R, const EnsGroup& G)
{
    const int* __restrict__ map = G.map[2];
    const int* __restrict__ off = G.off[2];
#pragma unroll 1
    for (int b = ens_first(G.base[2]); b < G.nblk[2]; b += gridDim.x) {
        VK_ENS_SETUP_B(b)
        p4_body<C, STAGE>(lb, nb, e.HN, e.Hpsi, e.K0, e.K1, e.K2, e.tw, e.g, ens_rk_args(e, STAGE));
        __syncthreads();
    }
}

__device__ __forceinline__ void ens_grid_barrier(const EnsStepArgs& A, int k)
{
    // this thread's stores visible device-wide (all XCDs' L2), then the
    // workgroup's arrival
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(A.bar + k, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = (long long)wall_clock64();
        while (__hip_atomic_load(A.bar + k, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x) {
            if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if ((long long)wall_clock64() - t0 > A.limit) {
                __hip_atomic_store(A.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__global__ void __launch_bounds__(256) VK_ENS_ROCC k_ens_step(const EnsReal* __restrict__ R, EnsStepArgs A)
{
    int k = 0;
#pragma unroll 1
    for (int stage = 1; stage <= 4; ++stage) {
#pragma unroll 1
        for (int g = 0; g < A.ngrp; ++g) VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p1<C>(R, A.grp[g])))
        ens_grid_barrier(A, k++);
#pragma unroll 1
        for (int g = 0; g < A.ngrp; ++g) VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p2<C>(R, A.grp[g])))
        ens_grid_barrier(A, k++);
#pragma unroll 1
        for (int g = 0; g < A.ngrp; ++g) VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p3<C>(R, A.grp[g])))
        ens_grid_barrier(A, k++);
#pragma unroll 1
        for (int g = 0; g < A.ngrp; ++g) {
            switch (stage) {
            case 1: VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p4<C, 1>(R, A.grp[g]))) break;
            case 2: VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p4<C, 2>(R, A.grp[g]))) break;
            case 3: VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p4<C, 3>(R, A.grp[g]))) break;
            default: VK_ENS_LKEYS(A.grp[g].key, (ens_loop_p4<C, 4>(R, A.grp[g]))) break;
            }
        }
        if (stage < 4) ens_grid_barrier(A, k++);
    }
}

// zero n counters (a kernel, so that a captured graph holds no memset node)
__global__ void k_zero_u32(unsigned* __restrict__ p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace vk
