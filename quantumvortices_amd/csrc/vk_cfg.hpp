// Per-size launch configuration and field layouts.
//
// Layouts in HBM (complex128 = 16 B):
//   x-space  X[y*N + x]                     (the reference's psi(i1=x,i2=y))
//   k-space  K[kx*N + ky]                   (the reference's kpsi(i2=ky,i1=kx),
//                                            FFTW_MPI_TRANSPOSED_OUT)
//   half-spectral H(y, kx) = DFT along x of a field, stored in panels of TC
//   adjacent kx columns:  H[(p*N + y)*TC + c],  kx = p*TC + c.
// A row pass (TR rows y per workgroup) reads/writes TR*TC*16 contiguous
// bytes per panel; a column pass (one panel = TC columns per workgroup)
// reads/writes one contiguous N*TC*16-byte block.
#pragma once
#include "vk_fft.hpp"

#ifndef VK_STAGGER
#define VK_STAGGER 32
#endif

namespace vk {

template <int N_, int S_, int TC_, int TR_, bool SPLIT_>
struct CfgT {
    static constexpr int N = N_;
    static constexpr int S = S_;          // points per thread per transform
    static constexpr int T = N_ / S_;     // threads per transform
    static constexpr int TC = TC_;        // columns per column-pass workgroup
    static constexpr int TR = TR_;        // rows per row-pass workgroup
    static constexpr int NP = N_ / TC_;   // panels
    static constexpr bool SPLIT = SPLIT_;
    static constexpr int LDSW = Lds<N_>::words * (SPLIT_ ? 1 : 2);  // doubles per transform
    static constexpr int COLT = TC_ * T;  // column-pass threads
    static constexpr int ROWT = TR_ * T;  // row-pass threads
    // LDS stride between the transforms of one workgroup: staggered by
    // VK_STAGGER/G doubles so that the G transforms interleaved across the
    // lanes of a wave start on disjoint bank pairs
    static constexpr int COLRS = LDSW + VK_STAGGER / TC_;
    static constexpr int ROWRS = LDSW + VK_STAGGER / TR_;
    static_assert(T % TC == 0, "TC must divide T");
    static_assert(T >= 1, "N/S >= 1");
};

template <int N> struct Cfg;
template <> struct Cfg<32> : CfgT<32, 8, 4, 4, true> {};
template <> struct Cfg<64> : CfgT<64, 8, 8, 8, true> {};
template <> struct Cfg<128> : CfgT<128, 8, 16, 16, true> {};
template <> struct Cfg<256> : CfgT<256, 16, 16, 16, true> {};
template <> struct Cfg<512> : CfgT<512, 16, 8, 8, true> {};
template <> struct Cfg<1024> : CfgT<1024, 16, 4, 4, true> {};
template <> struct Cfg<2048> : CfgT<2048, 16, 4, 4, true> {};
#ifndef VK4096_S
#define VK4096_S 16
#endif
// 4-column panels: the row passes gather 64-byte row pieces with one row
// per workgroup (P1 0.152 -> 0.139 ms, P3 0.186 -> 0.163 ms) and the column
// passes take two of a panel's columns per workgroup (P4 0.151 -> 0.170
// ms).  Per step, against 2-column panels on the same device: 2-3% faster
// without particles, within 1% with them (+0.9% and -0.4% on two
// devices).  The fields are bitwise the same: the layout moves data, the
// arithmetic is unchanged.
#ifndef VK4096_TC
#define VK4096_TC 4
#endif
#ifndef VK4096_TR
#define VK4096_TR 2
#endif
#ifndef VK4096_SPLIT
#define VK4096_SPLIT true
#endif
template <> struct Cfg<4096> : CfgT<4096, VK4096_S, VK4096_TC, VK4096_TR, VK4096_SPLIT> {};

// Shapes of the ensemble sizes, all with 256-thread workgroups
// (TC*T = TR*T = 256): variant 0 is Cfg<N>; variant 1 holds 8 points per
// thread instead of 16 (radix-8 stages, twice the threads per transform,
// half the panel width).  N = 128 has only variant 0 (16 points per thread
// would need 512-wide panels of 8 threads).
template <int N, int V> struct CfgE : Cfg<N> {};
template <> struct CfgE<256, 1> : CfgT<256, 8, 8, 8, true> {};
template <> struct CfgE<512, 1> : CfgT<512, 8, 4, 4, true> {};
template <> struct CfgE<1024, 1> : CfgT<1024, 8, 2, 2, true> {};
constexpr bool has_variant(int N, int V) { return V == 0 || (V == 1 && (N == 256 || N == 512 || N == 1024)); }

// The same shapes for pass bodies that run in a loop over blocks of several
// shapes (the persistent ensemble step): the thread index is read through an
// empty volatile asm, so that no value derived from it is hoisted out of the
// loop.  Hoisted, every shape's index arithmetic would stay live across the
// whole kernel and spill.
template <int N, int V> struct CfgL : CfgE<N, V> { static constexpr bool OPAQUE_TID = true; };

template <class C, class = void> struct opaque_tid { static constexpr bool value = false; };
template <class C> struct opaque_tid<C, decltype((void)C::OPAQUE_TID)> { static constexpr bool value = C::OPAQUE_TID; };

template <class C> __device__ __forceinline__ int tid_of()
{
    int u = threadIdx.x;
    if constexpr (opaque_tid<C>::value) {
        asm volatile("" : "+v"(u));
        __builtin_assume(u >= 0 && u < 1024);
    }
    return u;
}

// Runtime panel geometry.  Retained |kx| <= M = floor(N/3) (the 2/3 rule
// for 3 not dividing N).  Positive panels p < Pp cover kx in [0, M];
// negative panels p >= NP - Pn cover kx in [N-M, N-1].
struct Geo {
    int M;
    int Pp, Pn;
};

// Row blocks are handed to the 8 XCDs round-robin; remap block b of nb so
// that each XCD takes a contiguous run of row blocks (nb a multiple of 8).
__device__ __forceinline__ int xcd_remap_blk(int b, int nb)
{
    return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Signed wavenumber of index i: (i-1) - N*(i/(N/2+1)) with 1-based i
// (src/pspectral.f90:22); Nyquist maps to -N/2.
__device__ __forceinline__ int ksg(int i, int N) { return i < N / 2 ? i : i - N; }

// Dealias rule ksqr >= kalias**2 with kalias = N/3.0 (src/gpe.f90:346,
// src/mod_grid.f90:92).  For 3 not dividing N, N^2/9 is never within
// 1/9 of an integer, so the REAL comparison equals 9*ksqr >= N^2.
__device__ __forceinline__ bool kmasked(int kx, int ky, int N) { return 9 * (kx * kx + ky * ky) >= N * N; }

// panel pmode: 0 all, 1 retained both signs, 2 retained kx >= 0 only
__device__ __forceinline__ bool panel_stored(int p, int NP, const Geo& g, int pmode)
{
    if (pmode == 0) return true;
    if (pmode == 2) return p < g.Pp;
    return p < g.Pp || p >= NP - g.Pn;
}
__device__ __forceinline__ int panel_of_block(int b, int NP, const Geo& g, int pmode)
{
    if (pmode == 1 && b >= g.Pp) return NP - g.Pn + (b - g.Pp);
    return b;
}

template <class C>
__device__ __forceinline__ size_t hoff(int p, int y, int c)
{
    return ((size_t)p * C::N + y) * C::TC + c;
}

// Raw buffer access.  A buffer load whose byte offset is at or beyond the
// resource's size returns zero, and a buffer store there is dropped.  The
// passes issue each load unconditionally, with kOOB as the offset of the
// points they do not need, and select afterwards: a load inside a branch
// whose result is used in the same branch makes the compiler wait for that
// load before the next one is issued, which serialises the whole load
// phase.  Fields are at most 4096^2 complex doubles (256 MiB), so byte
// offsets fit in 32 bits and kOOB is beyond every field.
typedef __amdgpu_buffer_rsrc_t brsrc;
constexpr unsigned kOOB = 0x80000000u;
__device__ __forceinline__ brsrc mkbuf(const void* p, int N)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, N * N * 16, 0x00020000);
}
__device__ __forceinline__ cplx bld(brsrc r, unsigned off)
{
    const auto w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    cplx c;
    __builtin_memcpy(&c, &w, sizeof(c));
    return c;
}
__device__ __forceinline__ void bst(brsrc r, unsigned off, cplx c)
{
    __attribute__((ext_vector_type(4))) unsigned w;
    __builtin_memcpy(&w, &c, sizeof(c));
    __builtin_amdgcn_raw_buffer_store_b128(w, r, off, 0, 0);
}
// byte offset of element i, or kOOB
__device__ __forceinline__ unsigned boff(size_t i, bool use) { return use ? (unsigned)(i * 16) : kOOB; }

}  // namespace vk
