// HIP kernels of the direct J/K engine (gfx950, wave64).
//
// jk_kernel      one thread per (screened shell quartet, slice): Rys ERIs in
//                registers (FP64 or FP32 arithmetic), contracted with the
//                density at once (J and K fused), FP64 atomics into J/K.
//                Replaces getInts (src/qc.c:95-132) + formG (src/scf.f90:26-37).
// schwarz_kernel diagonal (ab|ab) blocks -> Schwarz bounds of every pair.
// block_kernel   explicit quartets written out (testing / store mode).
// prep / finish  density scaling + symmetrisation, J/K symmetrisation,
//                G = J - K/2.
#include <hip/hip_runtime.h>

#include "eri_rys.h"
#include "hfx_internal.h"

#define RYS_TABLE_QUALIFIER __device__ const
namespace hfx_dev {
#include "rys_tables.inc"
}
#undef RYS_TABLE_QUALIFIER

namespace {

__device__ HFX_INLINE RysTables dev_tables() {
    return RysTables{hfx_dev::RYS_CHEB, hfx_dev::RYS_HERM_U, hfx_dev::RYS_HERM_W};
}

struct DevAtomicAdd {
    __device__ HFX_INLINE void operator()(double* p, double v) const { unsafeAtomicAdd(p, v); }
};

template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void jk_kernel(JKArgs a) {
    constexpr bool SPA = split_a(LA, LB, LC, LD), SPC = split_c(LA, LB, LC, LD);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    const int t = blockIdx.x;
    const int row = a.tile_i0[t] + (int)(threadIdx.x / HFX_TILE_COLS);
    const int col = a.tile_j0[t] + (int)(threadIdx.x % HFX_TILE_COLS);
    if (row >= a.nbra) return;
    if (col < a.kbeg[row] || col >= a.kend[row]) return;
    const int split = blockIdx.y;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const ShellPair bra = a.pairs[a.bra_list[row]];
    const ShellPair ket = a.pairs[a.ket_list[col]];
    double s = 1.0;
    if (bra.A == bra.B) s *= 0.5;
    if (ket.A == ket.B) s *= 0.5;
    if (a.same && row == col) s *= 0.5;
    R eri[Q::NOUT];
    eri_quartet<LA, LB, LC, LD, SPA, SPC, R>(bra, ket, a.prims, sa, sc, dev_tables(), eri);
    digest_jk<LA, LB, LC, LD, SPA, SPC, R>(eri, bra.fa + sa, bra.fb, ket.fa + sc, ket.fb, a.D, a.n, s, a.J, a.K,
                                           DevAtomicAdd());
}

// q2bits[pair] = bits of max_{a,b} |(ab|ab)| (non-negative doubles order as integers)
template <int LA, int LB>
__global__ __launch_bounds__(256) void schwarz_kernel(const ShellPair* pairs, const PrimPair* prims, const int* list,
                                                      int npairs, unsigned long long* q2bits) {
    constexpr bool SPA = split_a(LA, LB, LA, LB), SPC = split_c(LA, LB, LA, LB);
    using Q = QClass<LA, LB, LA, LB, SPA, SPC>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npairs) return;
    const int split = blockIdx.y;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const int pid = list[i];
    const ShellPair pr = pairs[pid];
    double eri[Q::NOUT];
    eri_quartet<LA, LB, LA, LB, SPA, SPC, double>(pr, pr, prims, sa, sc, dev_tables(), eri);
    double m = 0.0;
#pragma unroll
    for (int a = 0; a < Q::NAt; a++)
#pragma unroll
        for (int b = 0; b < Q::NB; b++)
#pragma unroll
            for (int c = 0; c < Q::NCt; c++)
#pragma unroll
                for (int d = 0; d < Q::ND; d++) {
                    const int ga = SPA ? sa : a, gc = SPC ? sc : c;
                    if (ga == gc && b == d) m = fmax(m, fabs(eri[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d]));
                }
    if (m > 0.0) atomicMax(&q2bits[pid], (unsigned long long)__double_as_longlong(m));
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void block_kernel(BlockArgs a) {
    constexpr bool SPA = split_a(LA, LB, LC, LD), SPC = split_c(LA, LB, LC, LD);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    const int k = blockIdx.x;
    const int split = threadIdx.x;
    if (k >= a.nq || split >= nsplit(LA, LB, LC, LD)) return;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const int q = a.qidx[k];
    const ShellPair bra = a.pairs[a.bra[q]];
    const ShellPair ket = a.pairs[a.ket[q]];
    double eri[Q::NOUT];
    eri_quartet<LA, LB, LC, LD, SPA, SPC, double>(bra, ket, a.prims, sa, sc, dev_tables(), eri);
    double* o = a.out + a.off[q];
#pragma unroll
    for (int aa = 0; aa < Q::NAt; aa++)
#pragma unroll
        for (int b = 0; b < Q::NB; b++)
#pragma unroll
            for (int c = 0; c < Q::NCt; c++)
#pragma unroll
                for (int d = 0; d < Q::ND; d++) {
                    const int ga = SPA ? sa : aa, gc = SPC ? sc : c;
                    o[((ga * Q::NB + b) * Q::NC + gc) * Q::ND + d] = eri[((aa * Q::NB + b) * Q::NCt + c) * Q::ND + d];
                }
}

__global__ void prep_kernel(const double* __restrict__ P, const double* __restrict__ scale, double* __restrict__ D,
                            double* __restrict__ Jb, double* __restrict__ Kb, int n) {
    const size_t nn = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / n), j = (int)(t % n);
        D[t] = 0.5 * (P[(size_t)i * n + j] + P[(size_t)j * n + i]) * scale[i] * scale[j];
        Jb[t] = 0.0;
        Kb[t] = 0.0;
    }
}

__global__ void finish_kernel(const double* __restrict__ Jb, const double* __restrict__ Kb,
                              const double* __restrict__ scale, double* G, double* J, double* K, int n) {
    const size_t nn = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nn; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / n), j = (int)(t % n);
        const double sij = scale[i] * scale[j];
        const size_t tt = (size_t)j * n + i;
        const double jv = sij * (Jb[t] + Jb[tt]);
        const double kv = sij * (Kb[t] + Kb[tt]);
        if (G) G[t] = jv - 0.5 * kv;
        if (J) J[t] = jv;
        if (K) K[t] = kv;
    }
}

template <int LA, int LB, int LC, int LD, typename R>
hipError_t launch_jk_t(const JKArgs& a, int ntiles, hipStream_t s) {
    dim3 grid((unsigned)ntiles, (unsigned)nsplit(LA, LB, LC, LD));
    hipLaunchKernelGGL((jk_kernel<LA, LB, LC, LD, R>), grid, dim3(HFX_TILE_ROWS * HFX_TILE_COLS), 0, s, a);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_block_t(const BlockArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((block_kernel<LA, LB, LC, LD>), dim3((unsigned)a.nq), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <int X, int Y>
struct Combo {
    static constexpr int LA = pair_class_la(X), LB = pair_class_lb(X), LC = pair_class_la(Y), LD = pair_class_lb(Y);
    static hipError_t jk(bool fp32, const JKArgs& a, int nt, hipStream_t s) {
        return fp32 ? launch_jk_t<LA, LB, LC, LD, float>(a, nt, s) : launch_jk_t<LA, LB, LC, LD, double>(a, nt, s);
    }
    static hipError_t blk(const BlockArgs& a, hipStream_t s) { return launch_block_t<LA, LB, LC, LD>(a, s); }
};

typedef hipError_t (*jk_fn)(bool, const JKArgs&, int, hipStream_t);
typedef hipError_t (*blk_fn)(const BlockArgs&, hipStream_t);

#define HFX_ROW(X)                                                                                         \
    {Combo<X, 0>::jk, X >= 1 ? Combo<X, (X >= 1 ? 1 : 0)>::jk : nullptr,                                   \
     X >= 2 ? Combo<X, (X >= 2 ? 2 : 0)>::jk : nullptr, X >= 3 ? Combo<X, (X >= 3 ? 3 : 0)>::jk : nullptr,  \
     X >= 4 ? Combo<X, (X >= 4 ? 4 : 0)>::jk : nullptr, X >= 5 ? Combo<X, (X >= 5 ? 5 : 0)>::jk : nullptr}
#define HFX_BROW(X)                                                                                          \
    {Combo<X, 0>::blk, X >= 1 ? Combo<X, (X >= 1 ? 1 : 0)>::blk : nullptr,                                   \
     X >= 2 ? Combo<X, (X >= 2 ? 2 : 0)>::blk : nullptr, X >= 3 ? Combo<X, (X >= 3 ? 3 : 0)>::blk : nullptr,  \
     X >= 4 ? Combo<X, (X >= 4 ? 4 : 0)>::blk : nullptr, X >= 5 ? Combo<X, (X >= 5 ? 5 : 0)>::blk : nullptr}

const jk_fn JK_TABLE[6][6] = {HFX_ROW(0), HFX_ROW(1), HFX_ROW(2), HFX_ROW(3), HFX_ROW(4), HFX_ROW(5)};
const blk_fn BLK_TABLE[6][6] = {HFX_BROW(0), HFX_BROW(1), HFX_BROW(2), HFX_BROW(3), HFX_BROW(4), HFX_BROW(5)};

}  // namespace

int hfx_nsplit(int cb, int ck) {
    return nsplit(pair_class_la(cb), pair_class_lb(cb), pair_class_la(ck), pair_class_lb(ck));
}

hipError_t hfx_launch_jk(int cb, int ck, bool fp32, const JKArgs& a, int ntiles, hipStream_t s) {
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb || ntiles <= 0) return hipErrorInvalidValue;
    return JK_TABLE[cb][ck](fp32, a, ntiles, s);
}

hipError_t hfx_launch_block(int cb, int ck, const BlockArgs& a, hipStream_t s) {
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb || a.nq <= 0) return hipErrorInvalidValue;
    return BLK_TABLE[cb][ck](a, s);
}

hipError_t hfx_launch_schwarz(int cls, const ShellPair* pairs, const PrimPair* prims, const int* list, int npairs,
                              unsigned long long* q2bits, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((npairs + 255) / 256);
    switch (cls) {
        case 0:
            hipLaunchKernelGGL((schwarz_kernel<0, 0>), dim3(nb, hfx_nsplit(0, 0)), dim3(256), 0, s, pairs, prims,
                               list, npairs, q2bits);
            break;
        case 1:
            hipLaunchKernelGGL((schwarz_kernel<1, 0>), dim3(nb, hfx_nsplit(1, 1)), dim3(256), 0, s, pairs, prims,
                               list, npairs, q2bits);
            break;
        case 2:
            hipLaunchKernelGGL((schwarz_kernel<1, 1>), dim3(nb, hfx_nsplit(2, 2)), dim3(256), 0, s, pairs, prims,
                               list, npairs, q2bits);
            break;
        case 3:
            hipLaunchKernelGGL((schwarz_kernel<2, 0>), dim3(nb, hfx_nsplit(3, 3)), dim3(256), 0, s, pairs, prims,
                               list, npairs, q2bits);
            break;
        case 4:
            hipLaunchKernelGGL((schwarz_kernel<2, 1>), dim3(nb, hfx_nsplit(4, 4)), dim3(256), 0, s, pairs, prims,
                               list, npairs, q2bits);
            break;
        case 5:
            hipLaunchKernelGGL((schwarz_kernel<2, 2>), dim3(nb, hfx_nsplit(5, 5)), dim3(256), 0, s, pairs, prims,
                               list, npairs, q2bits);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t hfx_launch_prep(const double* P, const double* scale, double* D, double* Jb, double* Kb, int n,
                           hipStream_t s) {
    const size_t nn = (size_t)n * n;
    unsigned nb = (unsigned)((nn + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(prep_kernel, dim3(nb), dim3(256), 0, s, P, scale, D, Jb, Kb, n);
    return hipGetLastError();
}

hipError_t hfx_launch_finish(const double* Jb, const double* Kb, const double* scale, double* G, double* J, double* K,
                             int n, hipStream_t s) {
    const size_t nn = (size_t)n * n;
    unsigned nb = (unsigned)((nn + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(finish_kernel, dim3(nb), dim3(256), 0, s, Jb, Kb, scale, G, J, K, n);
    return hipGetLastError();
}
