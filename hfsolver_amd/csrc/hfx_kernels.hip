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

__device__ HFX_INLINE RysTables dev_tables(const double* boys = hfx_dev::BOYS_TAB) {
    return RysTables{hfx_dev::RYS_CHEB, hfx_dev::RYS_HERM_U, hfx_dev::RYS_HERM_W, boys};
}

__device__ HFX_INLINE double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One tile = HFX_TILE_ROWS bra rows x HFX_TILE_COLS ket columns; each
// wavefront owns one bra row (one bra shell pair), so J_ab is reduced across
// the wavefront before a single atomic per element.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void jk_kernel(JKArgs a) {
    constexpr bool SPA = split_a(LA, LB, LC, LD), SPC = split_c(LA, LB, LC, LD);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    // one- and two-root classes: Boys table staged in LDS for the closed-form rules
    const double* boys = hfx_dev::BOYS_TAB;
    if constexpr (Q::NR <= 2) {
        __shared__ double boys_lds[BOYS_SIZE];
        for (int k = threadIdx.x; k < BOYS_SIZE; k += blockDim.x) boys_lds[k] = hfx_dev::BOYS_TAB[k];
        __syncthreads();
        boys = boys_lds;
    }
    const int t = blockIdx.x;
    const int row = a.tile_i0[t] + (int)(threadIdx.x / HFX_TILE_COLS);
    const int col = a.tile_j0[t] + (int)(threadIdx.x % HFX_TILE_COLS);
    if (row >= a.nbra) return;  // uniform over the wavefront
    const bool active = col >= a.kbeg[row] && col < a.kend[row];
    const int split = blockIdx.y;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const int bp = a.bra_list[row];
    const ShellPair bra = a.pairs[bp];
    JKAcc<LA, LB, LC, LD, SPA, SPC, R> acc;
    jk_zero(acc);
    double s = 0.0;
    ShellPair ket;
    if (active) {
        const int kp = a.ket_list[col];
        ket = a.pairs[kp];
        s = 1.0;
        if (bra.A == bra.B) s *= 0.5;
        if (ket.A == ket.B) s *= 0.5;
        if (bp == kp) s *= 0.5;
        R eri[Q::NOUT];
        eri_quartet<LA, LB, LC, LD, SPA, SPC, R>(bra, ket, a.prims, sa, sc, dev_tables(boys), eri);
        jk_accumulate<LA, LB, LC, LD, SPA, SPC, R>(eri, bra.fa + sa, bra.fb, ket.fa + sc, ket.fb, a.D, a.n, acc);
    }
    // J_ab: same bra for the whole wavefront
    const bool lead = (threadIdx.x % 64) == 0;
    const double s2 = 2.0 * s;
    const int fa = bra.fa + sa, fb = bra.fb;
#pragma unroll
    for (int i = 0; i < Q::NAt; i++)
#pragma unroll
        for (int j = 0; j < Q::NB; j++) {
            const double v = wave_sum(s2 * double(acc.Jab[i][j]));
            if (lead && v != 0.0) unsafeAtomicAdd(&a.J[(size_t)(fa + i) * a.n + fb + j], v);
        }
    if (!active) return;
    const int fc = ket.fa + sc, fd = ket.fb, n = a.n;
#pragma unroll
    for (int c = 0; c < Q::NCt; c++)
#pragma unroll
        for (int d = 0; d < Q::ND; d++) unsafeAtomicAdd(&a.J[(size_t)(fc + c) * n + fd + d], s2 * double(acc.Jcd[c][d]));
#pragma unroll
    for (int i = 0; i < Q::NAt; i++) {
#pragma unroll
        for (int c = 0; c < Q::NCt; c++) unsafeAtomicAdd(&a.K[(size_t)(fa + i) * n + fc + c], s * double(acc.Kac[i][c]));
#pragma unroll
        for (int d = 0; d < Q::ND; d++) unsafeAtomicAdd(&a.K[(size_t)(fa + i) * n + fd + d], s * double(acc.Kad[i][d]));
    }
#pragma unroll
    for (int b = 0; b < Q::NB; b++) {
#pragma unroll
        for (int c = 0; c < Q::NCt; c++) unsafeAtomicAdd(&a.K[(size_t)(fb + b) * n + fc + c], s * double(acc.Kbc[b][c]));
#pragma unroll
        for (int d = 0; d < Q::ND; d++) unsafeAtomicAdd(&a.K[(size_t)(fb + b) * n + fd + d], s * double(acc.Kbd[b][d]));
    }
}

// Row mode: one block = one bra shell pair x a chunk of its kets (see RowArgs).
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void jk_row_kernel(RowArgs a) {
    constexpr bool SPA = split_a(LA, LB, LC, LD), SPC = split_c(LA, LB, LC, LD);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int NAt = Q::NAt, NB = Q::NB, NCt = Q::NCt, ND = Q::ND;
    extern __shared__ double smem[];
    const int ncols = a.ncols;
    const int nkl = (NAt + NB) * ncols;
    double* Kl = smem;
    double* Jl = smem + nkl;
    const double* boys = hfx_dev::BOYS_TAB;
    if constexpr (Q::NR <= 2) {
        if (a.boys_lds) {  // Boys table staged after the accumulators
            double* bl = smem + nkl + NAt * NB;
            for (int k = threadIdx.x; k < BOYS_SIZE; k += blockDim.x) bl[k] = hfx_dev::BOYS_TAB[k];
            boys = bl;
        }
    }
    for (int k = threadIdx.x; k < nkl + NAt * NB; k += blockDim.x) smem[k] = 0.0;
    __syncthreads();
    const int row = a.blk_row[blockIdx.x];
    const int split = blockIdx.y;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const int bp = a.bra_list[row];
    const ShellPair bra = a.pairs[bp];
    const int k0 = a.blk_k0[blockIdx.x];
    const int j0 = max(k0, a.kbeg[row]);
    const int j1 = min(k0 + a.chunk, a.kend[row]);
    const int fa = bra.fa + sa, fb = bra.fb, n = a.n;
    double Jab[NAt][NB];
#pragma unroll
    for (int i = 0; i < NAt; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) Jab[i][j] = 0.0;
    const double sbra = bra.A == bra.B ? 0.5 : 1.0;
    for (int col = j0 + (int)threadIdx.x; col < j1; col += (int)blockDim.x) {
        const int kp = a.ket_list[col];
        const ShellPair ket = a.pairs[kp];
        double s = sbra;
        if (ket.A == ket.B) s *= 0.5;
        if (bp == kp) s *= 0.5;
        R eri[Q::NOUT];
        eri_quartet<LA, LB, LC, LD, SPA, SPC, R>(bra, ket, a.prims, sa, sc, dev_tables(boys), eri);
        JKAcc<LA, LB, LC, LD, SPA, SPC, R> acc;
        jk_zero(acc);
        jk_accumulate<LA, LB, LC, LD, SPA, SPC, R>(eri, fa, fb, ket.fa + sc, ket.fb, a.D, n, acc);
        const double s2 = 2.0 * s;
#pragma unroll
        for (int i = 0; i < NAt; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) Jab[i][j] += s2 * double(acc.Jab[i][j]);
        const int fc = ket.fa + sc, fd = ket.fb;
#pragma unroll
        for (int c = 0; c < NCt; c++)
#pragma unroll
            for (int d = 0; d < ND; d++)
                unsafeAtomicAdd(&a.J[(size_t)(fc + c) * n + fd + d], s2 * double(acc.Jcd[c][d]));
        const int cc = ket.cpa + sc, cd = a.offd + ket.cpb;
#pragma unroll
        for (int i = 0; i < NAt; i++) {
#pragma unroll
            for (int c = 0; c < NCt; c++) atomicAdd(&Kl[i * ncols + cc + c], s * double(acc.Kac[i][c]));
#pragma unroll
            for (int d = 0; d < ND; d++) atomicAdd(&Kl[i * ncols + cd + d], s * double(acc.Kad[i][d]));
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
#pragma unroll
            for (int c = 0; c < NCt; c++) atomicAdd(&Kl[(NAt + b) * ncols + cc + c], s * double(acc.Kbc[b][c]));
#pragma unroll
            for (int d = 0; d < ND; d++) atomicAdd(&Kl[(NAt + b) * ncols + cd + d], s * double(acc.Kbd[b][d]));
        }
    }
    // J_ab: reduce over the wavefront, then over the block in LDS
#pragma unroll
    for (int i = 0; i < NAt; i++)
#pragma unroll
        for (int j = 0; j < NB; j++) {
            const double v = wave_sum(Jab[i][j]);
            if ((threadIdx.x % 64) == 0 && v != 0.0) atomicAdd(&Jl[i * NB + j], v);
        }
    __syncthreads();
    for (int k = threadIdx.x; k < nkl; k += blockDim.x) {
        const double v = Kl[k];
        if (v == 0.0) continue;
        const int r = k / ncols, col = k - r * ncols;
        const int rbf = r < NAt ? fa + r : fb + (r - NAt);
        const int cbf = (a.offd > 0 && col >= a.offd) ? a.bf_d[col - a.offd] : a.bf_c[col];
        unsafeAtomicAdd(&a.K[(size_t)rbf * n + cbf], v);
    }
    for (int k = threadIdx.x; k < NAt * NB; k += blockDim.x) {
        const double v = Jl[k];
        if (v != 0.0) unsafeAtomicAdd(&a.J[(size_t)(fa + k / NB) * n + fb + k % NB], v);
    }
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
constexpr int row_nab() {
    return (split_a(LA, LB, LC, LD) ? 1 : ncart(LA)) + ncart(LB);
}

template <int LA, int LB, int LC, int LD, typename R>
hipError_t launch_rows_t(const RowArgs& a, int nblocks, hipStream_t s) {
    constexpr int NAt = split_a(LA, LB, LC, LD) ? 1 : ncart(LA);
    const size_t lds = ((size_t)row_nab<LA, LB, LC, LD>() * a.ncols + NAt * ncart(LB) + (a.boys_lds ? BOYS_SIZE : 0)) *
                       sizeof(double);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)jk_row_kernel<LA, LB, LC, LD, R>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 32 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((unsigned)nblocks, (unsigned)nsplit(LA, LB, LC, LD));
    hipLaunchKernelGGL((jk_row_kernel<LA, LB, LC, LD, R>), grid, dim3(256), lds, s, a);
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
    static hipError_t rows(bool fp32, const RowArgs& a, int nb, hipStream_t s) {
        return fp32 ? launch_rows_t<LA, LB, LC, LD, float>(a, nb, s) : launch_rows_t<LA, LB, LC, LD, double>(a, nb, s);
    }
    static size_t row_lds(int ncols) {  // without the optional Boys table
        constexpr int NAt = split_a(LA, LB, LC, LD) ? 1 : ncart(LA);
        return ((size_t)row_nab<LA, LB, LC, LD>() * ncols + NAt * ncart(LB)) * sizeof(double);
    }
};

typedef hipError_t (*jk_fn)(bool, const JKArgs&, int, hipStream_t);
typedef hipError_t (*blk_fn)(const BlockArgs&, hipStream_t);
typedef hipError_t (*rows_fn)(bool, const RowArgs&, int, hipStream_t);
typedef size_t (*lds_fn)(int);

#define HFX_ROW(X)                                                                                         \
    {Combo<X, 0>::jk, X >= 1 ? Combo<X, (X >= 1 ? 1 : 0)>::jk : nullptr,                                   \
     X >= 2 ? Combo<X, (X >= 2 ? 2 : 0)>::jk : nullptr, X >= 3 ? Combo<X, (X >= 3 ? 3 : 0)>::jk : nullptr,  \
     X >= 4 ? Combo<X, (X >= 4 ? 4 : 0)>::jk : nullptr, X >= 5 ? Combo<X, (X >= 5 ? 5 : 0)>::jk : nullptr}
#define HFX_BROW(X)                                                                                          \
    {Combo<X, 0>::blk, X >= 1 ? Combo<X, (X >= 1 ? 1 : 0)>::blk : nullptr,                                   \
     X >= 2 ? Combo<X, (X >= 2 ? 2 : 0)>::blk : nullptr, X >= 3 ? Combo<X, (X >= 3 ? 3 : 0)>::blk : nullptr,  \
     X >= 4 ? Combo<X, (X >= 4 ? 4 : 0)>::blk : nullptr, X >= 5 ? Combo<X, (X >= 5 ? 5 : 0)>::blk : nullptr}

const jk_fn JK_TABLE[6][6] = {HFX_ROW(0), HFX_ROW(1), HFX_ROW(2), HFX_ROW(3), HFX_ROW(4), HFX_ROW(5)};

#define HFX_GROW(X, F)                                                                                       {Combo<X, 0>::F, X >= 1 ? Combo<X, (X >= 1 ? 1 : 0)>::F : nullptr,                                        X >= 2 ? Combo<X, (X >= 2 ? 2 : 0)>::F : nullptr, X >= 3 ? Combo<X, (X >= 3 ? 3 : 0)>::F : nullptr,       X >= 4 ? Combo<X, (X >= 4 ? 4 : 0)>::F : nullptr, X >= 5 ? Combo<X, (X >= 5 ? 5 : 0)>::F : nullptr}
const rows_fn ROWS_TABLE[6][6] = {HFX_GROW(0, rows), HFX_GROW(1, rows), HFX_GROW(2, rows),
                                  HFX_GROW(3, rows), HFX_GROW(4, rows), HFX_GROW(5, rows)};
const lds_fn LDS_TABLE[6][6] = {HFX_GROW(0, row_lds), HFX_GROW(1, row_lds), HFX_GROW(2, row_lds),
                                HFX_GROW(3, row_lds), HFX_GROW(4, row_lds), HFX_GROW(5, row_lds)};
const blk_fn BLK_TABLE[6][6] = {HFX_BROW(0), HFX_BROW(1), HFX_BROW(2), HFX_BROW(3), HFX_BROW(4), HFX_BROW(5)};

}  // namespace

int hfx_nsplit(int cb, int ck) {
    return nsplit(pair_class_la(cb), pair_class_lb(cb), pair_class_la(ck), pair_class_lb(ck));
}

hipError_t hfx_launch_jk(int cb, int ck, bool fp32, const JKArgs& a, int ntiles, hipStream_t s) {
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb || ntiles <= 0) return hipErrorInvalidValue;
    return JK_TABLE[cb][ck](fp32, a, ntiles, s);
}

size_t hfx_row_lds_bytes(int cb, int ck, int ncols) {
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb) return 0;
    return LDS_TABLE[cb][ck](ncols);
}

hipError_t hfx_launch_jk_rows(int cb, int ck, bool fp32, const RowArgs& a, int nblocks, hipStream_t s) {
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb || nblocks <= 0) return hipErrorInvalidValue;
    return ROWS_TABLE[cb][ck](fp32, a, nblocks, s);
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
