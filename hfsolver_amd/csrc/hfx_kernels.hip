// HIP kernels of the direct J/K engine (gfx950, wave64).
//
// jk_kernel      one thread per (screened shell quartet, slice): Rys ERIs in
//                registers (FP64 or FP32 arithmetic), contracted with the
//                density at once (J and K fused), FP64 atomics into J/K.
//                Replaces getInts (src/qc.c:95-132) + formG (src/scf.f90:26-37).
// schwarz_kernel diagonal (ab|ab) blocks -> Schwarz bounds of every pair.
// block_kernel   explicit quartets written out (testing / store mode).
// one_e_kernel   S, T, V (src/qc.c:6-93) and their nuclear gradient.
// prep / finish  density scaling + symmetrisation, J/K symmetrisation,
//                G = J - K/2.
#include <hip/hip_runtime.h>

#include <mutex>

#include "grad_rys.h"
#include "hfx_internal.h"
#include "one_e.h"

// The file is compiled once per part (-DHFX_PART=1..7) so the parts build in
// parallel: 1 tile-mode J/K, 2/3 row-mode J/K in FP64/FP32, 4/5 gradient in
// FP64/FP32, 6 Schwarz, explicit quartets, prep/finish and the dispatchers,
// 7 one-electron integrals and their gradient.
#ifndef HFX_PART
#error "compile with -DHFX_PART=1..7"
#endif

namespace {
#define RYS_TABLE_QUALIFIER __device__ const
namespace hfx_dev {
#include "rys_tables.inc"
}
#undef RYS_TABLE_QUALIFIER
}  // namespace

namespace {

__device__ HFX_INLINE RysTables dev_tables(const double* boys = hfx_dev::BOYS_TAB) {
    return RysTables{hfx_dev::RYS_POLY, hfx_dev::RYS_HERM_U, hfx_dev::RYS_HERM_W, boys, hfx_dev::BOYSF_TAB,
                     hfx_dev::RYS_CHEBF, hfx_dev::BOYS1_TAB};
}

// v of lane i - k within its row of 16 lanes (CTRL = row_shr:k), or of the
// broadcast lane of the previous row(s) (row_bcast:15 / :31); 0 where there is
// no source or the row is masked off. On the two 32-bit halves.
template <int CTRL, int ROW_MASK>
__device__ HFX_INLINE double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Sum over the wavefront, returned to every lane: inclusive prefix sums by DPP
// row shifts (1, 2, 4, 8) and row broadcasts (lane 15 -> next row, lane 31 ->
// rows 2, 3), all in registers; lane 63 holds the total.
__device__ HFX_INLINE double wave_sum(double v) {
    v += dpp_move<0x111, 0xf>(v);   // row_shr:1
    v += dpp_move<0x112, 0xf>(v);   // row_shr:2
    v += dpp_move<0x114, 0xf>(v);   // row_shr:4
    v += dpp_move<0x118, 0xf>(v);   // row_shr:8
    v += dpp_move<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

#if HFX_PART == 1
// One tile = HFX_TILE_ROWS bra rows x HFX_TILE_COLS ket columns; each
// wavefront owns one bra row (one bra shell pair), so J_ab is reduced across
// the wavefront before a single atomic per element.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void jk_kernel(JKArgs a) {
    constexpr bool F32 = sizeof(R) == 4;
    constexpr bool SPA = split_a(LA, LB, LC, LD, F32), SPC = split_c(LA, LB, LC, LD, F32);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    // the Boys table is read through the caches, as in the row-mode kernels
    const double* boys = hfx_dev::BOYS_TAB;
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
#endif

#if HFX_PART == 2 || HFX_PART == 3
// J_ab of one-row blocks summed in registers when the bra block has at most
// this many elements (otherwise reduced over the wavefront per quartet)
#ifndef HFX_JAB_REG_MAX
#define HFX_JAB_REG_MAX 36
#endif
// Row mode: one block = R bra shell pairs x a chunk of the ket table (see RowArgs);
// NARROW: the slicing of split_a_rows / split_c_rows for large molecules.
template <int LA, int LB, int LC, int LD, typename R, bool NARROW = false>
__global__ __launch_bounds__(hfx_row_threads(pair_class(LA, LB), pair_class(LC, LD), sizeof(R) == 4))
void jk_row_kernel(RowArgs a) {
    constexpr bool F32 = sizeof(R) == 4;
    constexpr bool SPA = split_a_rows(LA, LB, LC, LD, F32, NARROW), SPC = split_c_rows(LA, LB, LC, LD, F32, NARROW);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int NAt = Q::NAt, NB = Q::NB, NCt = Q::NCt, ND = Q::ND;
    constexpr int nrow = NAt + NB;             // LDS K rows per bra row
    extern __shared__ double smem[];
    __shared__ int rbf_s[HFX_ROWS_MAX * nrow];  // basis function of every LDS row
    const int ncols = a.ncols;
    const int row0 = hfx_ld_int(a.blk_row + blockIdx.x);
    const int nr = min(a.R, a.nbra - row0);    // bra rows of this block
    const int nkl = a.R * nrow * ncols;
    double* Kl = smem;                         // [R][nrow][ncols]
    double* Jl = smem + nkl;                   // [R][NAt][NB]
    R* Dl = reinterpret_cast<R*>(Jl + a.R * NAt * NB);   // [R][nrow][ncols] when a.dlds
    const int ndl = a.dlds ? nkl : 0;
    const int split = blockIdx.y;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const int n = a.n;
    if (threadIdx.x < nr * nrow) {
        const int r = threadIdx.x / nrow, i = threadIdx.x - r * nrow;
        const ShellPair bra = a.pairs[a.bra_list[row0 + r]];
        rbf_s[threadIdx.x] = i < NAt ? bra.fa + sa + i : bra.fb + (i - NAt);
    }
    for (int k = threadIdx.x; k < nkl + a.R * NAt * NB; k += blockDim.x) smem[k] = 0.0;
    __syncthreads();
    for (int k = threadIdx.x; k < nr * nrow * ncols && ndl > 0; k += blockDim.x) {
        const int ri = k / ncols, col = k - ri * ncols;
        const int cbf = (a.offd > 0 && col >= a.offd) ? a.bf_d[a.col0_d + col - a.offd] : a.bf_c[a.col0_c + col];
        Dl[k] = R(a.D[(size_t)rbf_s[ri] * n + cbf]);
    }
    __syncthreads();
    const KetTab& kt = a.kt;
    const int nk = kt.nk;
    const int k0 = hfx_ld_int(a.blk_k0 + blockIdx.x), k1 = k0 + a.chunk;
    const bool lead = (threadIdx.x % 64) == 0;
    using BP = typename BraPrimOf<R>::type;
    const BP* bprims;
    if constexpr (sizeof(R) == 4)
        bprims = a.primsf;
    else
        bprims = a.prims;
    // one-row blocks (the usual case): the bra's D_ab loaded once, its J_ab summed
    // in registers over all kets of the block and reduced over the wavefront at the end
    const bool onerow = nr == 1 && NAt * NB <= HFX_JAB_REG_MAX;
    double dab0[NAt][NB];   // wavefront-uniform: scalar loads
    double jab[NAt][NB];
    {
        const ShellPair bra0 = hfx_ld_pair(a.pairs + hfx_ld_int(a.bra_list + row0));
#pragma unroll
        for (int i = 0; i < NAt; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                dab0[i][j] = hfx_ld_dbl(a.D + (size_t)(bra0.fa + sa + i) * n + bra0.fb + j);
                jab[i][j] = 0.0;
            }
    }
    for (int s = 0; s < kt.S; s++) {
        // ket range of the block in sublist s: union of its rows' ranges within the chunk
        int lo = 1 << 30, hi = 0;
        for (int r = 0; r < nr; r++) {
            lo = min(lo, hfx_ld_int(a.kbeg + (row0 + r) * kt.S + s));
            hi = max(hi, hfx_ld_int(a.kend + (row0 + r) * kt.S + s));
        }
        const int j0 = max(k0, lo), j1 = min(k1, hi);
        if (j0 >= j1) continue;
        const int soff = kt.off[s], nks = kt.off[s + 1] - soff, npk = kt.np[s];
        // ket primitive planes in the arithmetic type of the kernel
        const R* sprim;
        if constexpr (sizeof(R) == 4)
            sprim = kt.primf + kt.psoff_f[s];
        else
            sprim = kt.prim + kt.psoff[s];
        // uniform trip count: every lane of a wavefront reaches the reductions
        for (int base = j0; base < j1; base += (int)blockDim.x) {
            const int col = base + (int)threadIdx.x;
            const bool incol = col < j1;
            KetInfo ki{0, 0, 0, 0};
            double cdx = 0, cdy = 0, cdz = 0;
            R Dcd[NCt][ND];
            if (incol) {
                ki = kt.info[col];
                cdx = kt.cd[col];
                cdy = kt.cd[nk + col];
                cdz = kt.cd[2 * nk + col];
#pragma unroll
                for (int c = 0; c < NCt; c++)
#pragma unroll
                    for (int d = 0; d < ND; d++) Dcd[c][d] = R(kt.Dcd[(size_t)((sc + c) * ND + d) * nk + col]);
            }
            // LDS columns: positions among the L_C (L_D) functions, less the window's first
            const int cc = (ki.cp & 0x7fff) - a.col0_c + sc;
            const int cd = a.offd + ((ki.cp >> 15) & 0x7fff) - (a.offd > 0 ? a.col0_d : a.col0_c);
            const double sket = (ki.cp >> 30) & 1 ? 0.5 : 1.0;
            const KetPrimsSoA<R> kp{sprim, col - soff, npk, ket_pitch(nks)};
            R Jcd[NCt][ND];
#pragma unroll
            for (int c = 0; c < NCt; c++)
#pragma unroll
                for (int d = 0; d < ND; d++) Jcd[c][d] = R(0);
            for (int r = 0; r < nr; r++) {
                const int row = row0 + r;
                const int bp = hfx_ld_int(a.bra_list + row);
                const ShellPair bra = hfx_ld_pair(a.pairs + bp);
                bool act = incol && col >= hfx_ld_int(a.kbeg + row * kt.S + s) &&
                           col < hfx_ld_int(a.kend + row * kt.S + s);
                if (a.dsh != nullptr && act) {
                    // density-weighted screening: the largest density factor of the quartet
                    const int shk = kt.shcd[col], shc = shk & 0xffff, shd = shk >> 16;
                    const float* dA = a.dsh + (size_t)bra.A * a.nsh;
                    const float* dB = a.dsh + (size_t)bra.B * a.nsh;
                    const float m = fmaxf(fmaxf(fmaxf(dA[shc], dA[shd]), fmaxf(dB[shc], dB[shd])),
                                          fmaxf(dA[bra.B], a.dsh[(size_t)shc * a.nsh + shd]));
                    if ((float)bra.Q * kt.qk[col] * m < a.dthr) act = false;
                }
                const int fa = bra.fa + sa, fb = bra.fb;
                JKAcc<LA, LB, LC, LD, SPA, SPC, R> acc;
                jk_zero(acc);
                double sf = 0.0;
                if (act) {
                    sf = bra.A == bra.B ? 0.5 * sket : sket;
                    if (bp == ki.pid) sf *= 0.5;
                    R eri[Q::NOUT];
                    eri_quartet_k<LA, LB, LC, LD, SPA, SPC, R>(bra, bprims, cdx, cdy, cdz, kp, sa, sc,
                                                               dev_tables(), eri);
                    JKDens<LA, LB, LC, LD, SPA, SPC, R> dn;
                    if (onerow) {
#pragma unroll
                        for (int i = 0; i < NAt; i++)
#pragma unroll
                            for (int j = 0; j < NB; j++) dn.ab[i][j] = R(dab0[i][j]);
                    } else {
#pragma unroll
                        for (int i = 0; i < NAt; i++)
#pragma unroll
                            for (int j = 0; j < NB; j++) dn.ab[i][j] = R(a.D[(size_t)(fa + i) * n + fb + j]);
                    }
#pragma unroll
                    for (int c = 0; c < NCt; c++)
#pragma unroll
                        for (int d = 0; d < ND; d++) dn.cd[c][d] = Dcd[c][d];
                    if (ndl > 0) {
                        const R* Dr = Dl + r * nrow * ncols;
#pragma unroll
                        for (int i = 0; i < NAt; i++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) dn.ac[i][c] = Dr[i * ncols + cc + c];
#pragma unroll
                            for (int d = 0; d < ND; d++) dn.ad[i][d] = Dr[i * ncols + cd + d];
                        }
#pragma unroll
                        for (int b = 0; b < NB; b++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) dn.bc[b][c] = Dr[(NAt + b) * ncols + cc + c];
#pragma unroll
                            for (int d = 0; d < ND; d++) dn.bd[b][d] = Dr[(NAt + b) * ncols + cd + d];
                        }
                    } else {
                        const int fc = ki.fa + sc, fd = ki.fb;
#pragma unroll
                        for (int i = 0; i < NAt; i++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) dn.ac[i][c] = R(a.D[(size_t)(fa + i) * n + fc + c]);
#pragma unroll
                            for (int d = 0; d < ND; d++) dn.ad[i][d] = R(a.D[(size_t)(fa + i) * n + fd + d]);
                        }
#pragma unroll
                        for (int b = 0; b < NB; b++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) dn.bc[b][c] = R(a.D[(size_t)(fb + b) * n + fc + c]);
#pragma unroll
                            for (int d = 0; d < ND; d++) dn.bd[b][d] = R(a.D[(size_t)(fb + b) * n + fd + d]);
                        }
                    }
                    jk_contract<LA, LB, LC, LD, SPA, SPC, R>(eri, dn, acc);
                    const R rs = R(sf);
#pragma unroll
                    for (int c = 0; c < NCt; c++)
#pragma unroll
                        for (int d = 0; d < ND; d++) Jcd[c][d] += rs * acc.Jcd[c][d];
                    double* Kr = Kl + r * nrow * ncols;
#pragma unroll
                    for (int i = 0; i < NAt; i++) {
#pragma unroll
                        for (int c = 0; c < NCt; c++) atomicAdd(&Kr[i * ncols + cc + c], sf * double(acc.Kac[i][c]));
#pragma unroll
                        for (int d = 0; d < ND; d++) atomicAdd(&Kr[i * ncols + cd + d], sf * double(acc.Kad[i][d]));
                    }
#pragma unroll
                    for (int b = 0; b < NB; b++) {
#pragma unroll
                        for (int c = 0; c < NCt; c++)
                            atomicAdd(&Kr[(NAt + b) * ncols + cc + c], sf * double(acc.Kbc[b][c]));
#pragma unroll
                        for (int d = 0; d < ND; d++)
                            atomicAdd(&Kr[(NAt + b) * ncols + cd + d], sf * double(acc.Kbd[b][d]));
                    }
                }
                if (onerow) {
#pragma unroll
                    for (int i = 0; i < NAt; i++)
#pragma unroll
                        for (int j = 0; j < NB; j++) jab[i][j] += 2.0 * sf * double(acc.Jab[i][j]);
                } else {
                    // J_ab of row r: reduce over the wavefront, one LDS add per wavefront
#pragma unroll
                    for (int i = 0; i < NAt; i++)
#pragma unroll
                        for (int j = 0; j < NB; j++) {
                            const double v = wave_sum(2.0 * sf * double(acc.Jab[i][j]));
                            if (lead && v != 0.0) atomicAdd(&Jl[(r * NAt + i) * NB + j], v);
                        }
                }
            }
            if (incol) {
#pragma unroll
                for (int c = 0; c < NCt; c++)
#pragma unroll
                    for (int d = 0; d < ND; d++) {
                        const double v = 2.0 * double(Jcd[c][d]);
                        if (v != 0.0) unsafeAtomicAdd(&kt.Jcd[(size_t)((sc + c) * ND + d) * nk + col], v);
                    }
            }
        }
    }
    if (onerow) {
#pragma unroll
        for (int i = 0; i < NAt; i++)
#pragma unroll
            for (int j = 0; j < NB; j++) {
                const double v = wave_sum(jab[i][j]);
                if (lead && v != 0.0) atomicAdd(&Jl[i * NB + j], v);
            }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nr * nrow * ncols; k += blockDim.x) {
        const double v = Kl[k];
        if (v == 0.0) continue;
        const int ri = k / ncols, col = k - ri * ncols;
        const int cbf = (a.offd > 0 && col >= a.offd) ? a.bf_d[a.col0_d + col - a.offd] : a.bf_c[a.col0_c + col];
        unsafeAtomicAdd(&a.K[(size_t)rbf_s[ri] * n + cbf], v);
    }
    for (int k = threadIdx.x; k < nr * NAt * NB; k += blockDim.x) {
        const double v = Jl[k];
        if (v == 0.0) continue;
        const int r = k / (NAt * NB), ij = k - r * NAt * NB;
        unsafeAtomicAdd(&a.J[(size_t)rbf_s[r * nrow + ij / NB] * n + rbf_s[r * nrow + NAt + ij % NB]], v);
    }
}

#endif

#if HFX_PART == 4 || HFX_PART == 5
// Two-electron gradient, row mode: the blocks and ket tables of the row-mode
// J/K launch (see GradRowArgs). Per quartet gam = s (4 D_ab D_cd - D_ac D_bd -
// D_ad D_bc); A and B components reduced over the wavefront per row, C and D
// (= -(A + B + C), translational invariance) by LDS atomics on their atoms.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(hfx_grad_threads(pair_class(LA, LB), pair_class(LC, LD), sizeof(R) == 4))
void grad_row_kernel(GradRowArgs a) {
    constexpr bool G32 = sizeof(R) == 4;
    constexpr bool SPA = gsplit_a(LA, LB, LC, LD, G32), SPC = gsplit_c(LA, LB, LC, LD, G32);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int NAt = Q::NAt, NB = Q::NB, NCt = Q::NCt, ND = Q::ND;
    constexpr int nrow = NAt + NB;
    extern __shared__ double smem[];
    __shared__ int rbf_s[HFX_ROWS_MAX * nrow];
    const int ncols = a.ncols;
    const int row0 = hfx_ld_int(a.blk_row + blockIdx.x);
    const int nr = min(a.R, a.nbra - row0);
    const int ndl = a.dlds ? a.R * nrow * ncols : 0;
    R* Dl = reinterpret_cast<R*>(smem);
    double* gl = smem + (sizeof(R) == 8 ? ndl : (ndl + 1) / 2);   // 3 * nat
    const int split = blockIdx.y;
    const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
    const int sc = SPC ? split % Q::NC : 0;
    const int n = a.n;
    if (threadIdx.x < nr * nrow) {
        const int r = threadIdx.x / nrow, i = threadIdx.x - r * nrow;
        const ShellPair bra = a.pairs[a.bra_list[row0 + r]];
        rbf_s[threadIdx.x] = i < NAt ? bra.fa + sa + i : bra.fb + (i - NAt);
    }
    for (int k = threadIdx.x; k < 3 * a.nat; k += blockDim.x) gl[k] = 0.0;
    __syncthreads();
    for (int k = threadIdx.x; k < nr * nrow * ncols && ndl > 0; k += blockDim.x) {
        const int ri = k / ncols, col = k - ri * ncols;
        const int cbf = (a.offd > 0 && col >= a.offd) ? a.bf_d[a.col0_d + col - a.offd] : a.bf_c[a.col0_c + col];
        Dl[k] = R(a.D[(size_t)rbf_s[ri] * n + cbf]);
    }
    __syncthreads();
    const KetTab& kt = a.kt;
    const int nk = kt.nk;
    const int k0 = hfx_ld_int(a.blk_k0 + blockIdx.x), k1 = k0 + a.chunk;
    const bool lead = (threadIdx.x % 64) == 0;
    for (int s = 0; s < kt.S; s++) {
        int lo = 1 << 30, hi = 0;
        for (int r = 0; r < nr; r++) {
            lo = min(lo, hfx_ld_int(a.kbeg + (row0 + r) * kt.S + s));
            hi = max(hi, hfx_ld_int(a.kend + (row0 + r) * kt.S + s));
        }
        const int j0 = max(k0, lo), j1 = min(k1, hi);
        if (j0 >= j1) continue;
        const int soff = kt.off[s], nks = kt.off[s + 1] - soff, npk = kt.np[s];
        const double* sprim = kt.prim + kt.psoff[s];
        for (int base = j0; base < j1; base += (int)blockDim.x) {
            const int col = base + (int)threadIdx.x;
            const bool incol = col < j1;
            KetInfo ki{0, 0, 0, 0};
            double cdx = 0, cdy = 0, cdz = 0;
            R Dcd[NCt][ND];
            int atc = 0, atd = 0;
            if (incol) {
                ki = kt.info[col];
                cdx = kt.cd[col];
                cdy = kt.cd[nk + col];
                cdz = kt.cd[2 * nk + col];
#pragma unroll
                for (int c = 0; c < NCt; c++)
#pragma unroll
                    for (int d = 0; d < ND; d++) Dcd[c][d] = R(kt.Dcd[(size_t)((sc + c) * ND + d) * nk + col]);
                atc = a.bf_atom[ki.fa];
                atd = a.bf_atom[ki.fb];
            }
            // LDS columns: positions among the L_C (L_D) functions, less the window's first
            const int cc = (ki.cp & 0x7fff) - a.col0_c + sc;
            const int cd = a.offd + ((ki.cp >> 15) & 0x7fff) - (a.offd > 0 ? a.col0_d : a.col0_c);
            const double sket = (ki.cp >> 30) & 1 ? 0.5 : 1.0;
            const KetPrimsSoA<double> kp{sprim, col - soff, npk, ket_pitch(nks)};
            double gcd[6] = {0, 0, 0, 0, 0, 0};   // C, D components summed over the rows
            for (int r = 0; r < nr; r++) {
                const int row = row0 + r;
                const int bp = hfx_ld_int(a.bra_list + row);
                const ShellPair bra = hfx_ld_pair(a.pairs + bp);
                const bool act = incol && col >= hfx_ld_int(a.kbeg + row * kt.S + s) &&
                                 col < hfx_ld_int(a.kend + row * kt.S + s);
                const int fa = bra.fa + sa, fb = bra.fb;
                double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                if (act) {
                    double sf = bra.A == bra.B ? 0.5 * sket : sket;
                    if (bp == ki.pid) sf *= 0.5;
                    R Dab[NAt][NB], Dac[NAt][NCt], Dad[NAt][ND], Dbc[NB][NCt], Dbd[NB][ND];
#pragma unroll
                    for (int i = 0; i < NAt; i++)
#pragma unroll
                        for (int j = 0; j < NB; j++) Dab[i][j] = R(a.D[(size_t)(fa + i) * n + fb + j]);
                    if (ndl > 0) {
                        const R* Dr = Dl + r * nrow * ncols;
#pragma unroll
                        for (int i = 0; i < NAt; i++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) Dac[i][c] = Dr[i * ncols + cc + c];
#pragma unroll
                            for (int d = 0; d < ND; d++) Dad[i][d] = Dr[i * ncols + cd + d];
                        }
#pragma unroll
                        for (int b = 0; b < NB; b++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) Dbc[b][c] = Dr[(NAt + b) * ncols + cc + c];
#pragma unroll
                            for (int d = 0; d < ND; d++) Dbd[b][d] = Dr[(NAt + b) * ncols + cd + d];
                        }
                    } else {
                        const int fc = ki.fa + sc, fd = ki.fb;
#pragma unroll
                        for (int i = 0; i < NAt; i++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) Dac[i][c] = R(a.D[(size_t)(fa + i) * n + fc + c]);
#pragma unroll
                            for (int d = 0; d < ND; d++) Dad[i][d] = R(a.D[(size_t)(fa + i) * n + fd + d]);
                        }
#pragma unroll
                        for (int b = 0; b < NB; b++) {
#pragma unroll
                            for (int c = 0; c < NCt; c++) Dbc[b][c] = R(a.D[(size_t)(fb + b) * n + fc + c]);
#pragma unroll
                            for (int d = 0; d < ND; d++) Dbd[b][d] = R(a.D[(size_t)(fb + b) * n + fd + d]);
                        }
                    }
                    R gam[Q::NOUT];
                    const R rs = R(sf);
#pragma unroll
                    for (int i = 0; i < NAt; i++)
#pragma unroll
                        for (int b = 0; b < NB; b++)
#pragma unroll
                            for (int c = 0; c < NCt; c++)
#pragma unroll
                                for (int d = 0; d < ND; d++)
                                    gam[((i * NB + b) * NCt + c) * ND + d] =
                                        rs * (R(4) * Dab[i][b] * Dcd[c][d] - Dac[i][c] * Dbd[b][d] - Dad[i][d] * Dbc[b][c]);
                    grad_quartet_k<LA, LB, LC, LD, SPA, SPC, R>(bra, a.prims, a.pexp, cdx, cdy, cdz, kp, sa, sc,
                                                                dev_tables(), gam, g);
#pragma unroll
                    for (int x = 0; x < 3; x++) {
                        gcd[x] += g[6 + x];
                        gcd[3 + x] -= g[x] + g[3 + x] + g[6 + x];
                    }
                }
                // A and B of row r: reduce over the wavefront, one LDS add per wavefront
                const int ata = a.bf_atom[bra.fa], atb = a.bf_atom[bra.fb];
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    const double v = wave_sum(g[k]);
                    if (lead && v != 0.0) atomicAdd(&gl[3 * (k < 3 ? ata : atb) + k % 3], v);
                }
            }
            if (incol) {
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    if (gcd[x] != 0.0) atomicAdd(&gl[3 * atc + x], gcd[x]);
                    if (gcd[3 + x] != 0.0) atomicAdd(&gl[3 * atd + x], gcd[3 + x]);
                }
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * a.nat; k += blockDim.x)
        if (gl[k] != 0.0) unsafeAtomicAdd(&a.grad[k], gl[k]);
}

// Two-electron gradient: one block = one bra shell pair x a chunk of its kets.
template <int LA, int LB, int LC, int LD, typename R>
__global__ __launch_bounds__(256) void grad_kernel(GradArgs a) {
    constexpr bool G32 = sizeof(R) == 4;
    constexpr bool SPA = gsplit_a(LA, LB, LC, LD, G32), SPC = gsplit_c(LA, LB, LC, LD, G32);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    extern __shared__ double smem[];
    double* gl = smem;  // 3 * nat
    // the Boys table is read through the caches, as in the row-mode kernels
    const double* boys = hfx_dev::BOYS_TAB;
    for (int k = threadIdx.x; k < 3 * a.nat; k += blockDim.x) gl[k] = 0.0;
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
    const double sbra = bra.A == bra.B ? 0.5 : 1.0;
    double gab[6] = {0, 0, 0, 0, 0, 0};
    for (int col = j0 + (int)threadIdx.x; col < j1; col += (int)blockDim.x) {
        const int kp = a.ket_list[col];
        const ShellPair ket = a.pairs[kp];
        double s = sbra;
        if (ket.A == ket.B) s *= 0.5;
        if (bp == kp) s *= 0.5;
        R gam[Q::NOUT];
        grad_gamma<LA, LB, LC, LD, SPA, SPC, R>(bra.fa + sa, bra.fb, ket.fa + sc, ket.fb, a.D, a.n, s, gam);
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        grad_quartet<LA, LB, LC, LD, SPA, SPC, R>(bra, ket, a.prims, a.pexp, sa, sc, dev_tables(boys), gam, g);
#pragma unroll
        for (int k = 0; k < 6; k++) gab[k] += g[k];
        const int atc = a.shell_atom[ket.A], atd = a.shell_atom[ket.B];
#pragma unroll
        for (int x = 0; x < 3; x++) {
            atomicAdd(&gl[3 * atc + x], g[6 + x]);
            atomicAdd(&gl[3 * atd + x], -(g[x] + g[3 + x] + g[6 + x]));
        }
    }
    const int ata = a.shell_atom[bra.A], atb = a.shell_atom[bra.B];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const double v = wave_sum(gab[k]);
        if ((threadIdx.x % 64) == 0 && v != 0.0) atomicAdd(&gl[3 * (k < 3 ? ata : atb) + k % 3], v);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * a.nat; k += blockDim.x)
        if (gl[k] != 0.0) unsafeAtomicAdd(&a.grad[k], gl[k]);
}
#endif

#if HFX_PART == 6
// q2bits[pair] = bits of max_{a,b} |(ab|ab)| over the basis functions a, b
// (non-negative doubles order as integers). The kernels evaluate integrals
// with the shell's coefficients; a function's own integrals carry its scale
// (scale[f], e.g. sqrt(3) for d_xy against d_xx), so (ab|ab) of the functions
// is scale_a^2 scale_b^2 times the shell-coefficient value.
template <int LA, int LB>
__global__ __launch_bounds__(256) void schwarz_kernel(const ShellPair* pairs, const PrimPair* prims, const int* list,
                                                      int npairs, const double* scale,
                                                      unsigned long long* q2bits) {
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
                    if (ga == gc && b == d) {
                        const double s2 = scale[pr.fa + ga] * scale[pr.fb + b];
                        m = fmax(m, s2 * s2 * fabs(eri[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d]));
                    }
                }
    if (m > 0.0) atomicMax(&q2bits[pid], (unsigned long long)__double_as_longlong(m));
}

// Position of (ij|kl) in the 8-fold packed array: ij = i(i+1)/2 + j for
// i >= j, the same for kl, then the larger pair index first (64-bit).
__device__ HFX_INLINE long long packed_index(long long i, long long j, long long k, long long l) {
    const long long ij = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
    const long long kl = k >= l ? k * (k + 1) / 2 + l : l * (l + 1) / 2 + k;
    return ij >= kl ? ij * (ij + 1) / 2 + kl : kl * (kl + 1) / 2 + ij;
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
    if (a.packed) {
#pragma unroll
        for (int aa = 0; aa < Q::NAt; aa++)
#pragma unroll
            for (int b = 0; b < Q::NB; b++)
#pragma unroll
                for (int c = 0; c < Q::NCt; c++)
#pragma unroll
                    for (int d = 0; d < Q::ND; d++) {
                        const int i = bra.fa + (SPA ? sa : aa), j = bra.fb + b;
                        const int k2 = ket.fa + (SPC ? sc : c), l = ket.fb + d;
                        const double v = eri[((aa * Q::NB + b) * Q::NCt + c) * Q::ND + d];
                        a.out[packed_index(i, j, k2, l)] = v * a.scale[i] * a.scale[j] * a.scale[k2] * a.scale[l];
                    }
        return;
    }
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

// D_cd of every ket of a table (field-major) and zeroed J_cd accumulators.
__global__ void ket_gather_kernel(KetTab kt, int ncomp, int nd, const double* __restrict__ D, int n) {
    const size_t tot = (size_t)ncomp * kt.nk;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
        const int comp = (int)(idx / kt.nk), k = (int)(idx - (size_t)comp * kt.nk);
        const KetInfo ki = kt.info[k];
        const int c = comp / nd, d = comp - c * nd;
        kt.Dcd[idx] = D[(size_t)(ki.fa + c) * n + ki.fb + d];
        kt.Jcd[idx] = 0.0;
    }
}

// Jb += J_cd of every ket (each ket is a distinct shell pair, so every element
// has one writer).
__global__ void ket_scatter_kernel(KetTab kt, int ncomp, int nd, double* __restrict__ Jb, int n) {
    const size_t tot = (size_t)ncomp * kt.nk;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
        const int comp = (int)(idx / kt.nk), k = (int)(idx - (size_t)comp * kt.nk);
        const KetInfo ki = kt.info[k];
        const int c = comp / nd, d = comp - c * nd;
        Jb[(size_t)(ki.fa + c) * n + ki.fb + d] += kt.Jcd[idx];
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
#endif

// table of a function template F<X, Y> over class combinations X >= Y
#define HFX_TROW(F, X)                                                                                    \
    {F<X, 0>, X >= 1 ? F<X, (X >= 1 ? 1 : 0)> : nullptr, X >= 2 ? F<X, (X >= 2 ? 2 : 0)> : nullptr,        \
     X >= 3 ? F<X, (X >= 3 ? 3 : 0)> : nullptr, X >= 4 ? F<X, (X >= 4 ? 4 : 0)> : nullptr,                 \
     X >= 5 ? F<X, (X >= 5 ? 5 : 0)> : nullptr}
#define HFX_TABLE(F) {HFX_TROW(F, 0), HFX_TROW(F, 1), HFX_TROW(F, 2), HFX_TROW(F, 3), HFX_TROW(F, 4), HFX_TROW(F, 5)}
#define HFX_CLS(X) pair_class_la(X), pair_class_lb(X)

#if HFX_PART == 1
template <int X, int Y>
hipError_t jk_xy(bool fp32, const JKArgs& a, int ntiles, hipStream_t s) {
    dim3 grid((unsigned)ntiles, (unsigned)nsplit(HFX_CLS(X), HFX_CLS(Y), fp32));
    if (fp32)
        hipLaunchKernelGGL((jk_kernel<HFX_CLS(X), HFX_CLS(Y), float>), grid, dim3(HFX_TILE_ROWS * HFX_TILE_COLS), 0, s, a);
    else
        hipLaunchKernelGGL((jk_kernel<HFX_CLS(X), HFX_CLS(Y), double>), grid, dim3(HFX_TILE_ROWS * HFX_TILE_COLS), 0, s,
                           a);
    return hipGetLastError();
}
typedef hipError_t (*jk_fn)(bool, const JKArgs&, int, hipStream_t);
const jk_fn JK_TABLE[6][6] = HFX_TABLE(jk_xy);
#endif

#if HFX_PART == 2 || HFX_PART == 3
typedef HFX_REAL_T RowReal;
template <int X, int Y, bool NARROW = false>
hipError_t rows_xy(const RowArgs& a, int nblocks, hipStream_t s) {
    constexpr bool F32 = sizeof(RowReal) == 4;
    constexpr int NAt = split_a_rows(HFX_CLS(X), HFX_CLS(Y), F32, NARROW) ? 1 : ncart(pair_class_la(X));
    constexpr int NB = ncart(pair_class_lb(X));
    const size_t lds = hfx_row_lds(NAt + NB, NAt * NB, a.ncols, a.R, a.dlds, F32);
    if (lds > HFX_ROW_LDS_MAX || a.R < 1 || a.R > HFX_ROWS_MAX || a.kt.S < 1 || a.kt.S > HFX_SUB_MAX)
        return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)jk_row_kernel<HFX_CLS(X), HFX_CLS(Y), RowReal, NARROW>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, HFX_ROW_LDS_MAX);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((unsigned)nblocks, (unsigned)nsplit_rows(HFX_CLS(X), HFX_CLS(Y), F32, NARROW));
    hipLaunchKernelGGL((jk_row_kernel<HFX_CLS(X), HFX_CLS(Y), RowReal, NARROW>), grid,
                       dim3(hfx_row_threads(X, Y, F32)), lds, s, a);
    return hipGetLastError();
}
// the narrow form where it differs (no second instantiation elsewhere)
template <int X, int Y>
hipError_t rows_narrow_xy(const RowArgs& a, int nblocks, hipStream_t s) {
    if constexpr (narrow_differs(HFX_CLS(X), HFX_CLS(Y), sizeof(RowReal) == 4))
        return rows_xy<X, Y, true>(a, nblocks, s);
    else
        return rows_xy<X, Y, false>(a, nblocks, s);
}
typedef hipError_t (*rows_fn)(const RowArgs&, int, hipStream_t);
const rows_fn ROWS_TABLE[6][6] = HFX_TABLE(rows_xy);
const rows_fn ROWS_NARROW_TABLE[6][6] = HFX_TABLE(rows_narrow_xy);
#endif

#if HFX_PART == 4 || HFX_PART == 5
typedef HFX_REAL_T GradReal;
template <int X, int Y>
hipError_t grad_xy(const GradArgs& a, int nblocks, hipStream_t s) {
    const size_t lds = 3 * (size_t)a.nat * sizeof(double);
    if (lds > HFX_GRAD_LDS_MAX) return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)grad_kernel<HFX_CLS(X), HFX_CLS(Y), GradReal>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)HFX_GRAD_LDS_MAX);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((unsigned)nblocks, (unsigned)gnsplit(HFX_CLS(X), HFX_CLS(Y), sizeof(GradReal) == 4));
    hipLaunchKernelGGL((grad_kernel<HFX_CLS(X), HFX_CLS(Y), GradReal>), grid, dim3(256), lds, s, a);
    return hipGetLastError();
}
typedef hipError_t (*grad_fn)(const GradArgs&, int, hipStream_t);
const grad_fn GRAD_TABLE[6][6] = HFX_TABLE(grad_xy);

template <int X, int Y>
hipError_t grad_rows_xy(const GradRowArgs& a, int nblocks, hipStream_t s) {
    constexpr int NAt = gsplit_a(HFX_CLS(X), HFX_CLS(Y), sizeof(GradReal) == 4) ? 1 : ncart(pair_class_la(X));
    constexpr int NB = ncart(pair_class_lb(X));
    const size_t lds = hfx_grad_row_lds(NAt + NB, a.ncols, a.R, a.dlds, sizeof(GradReal) == 4, a.nat);
    if (lds > HFX_ROW_LDS_MAX || a.R < 1 || a.R > HFX_ROWS_MAX || a.kt.S < 1 || a.kt.S > HFX_SUB_MAX)
        return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)grad_row_kernel<HFX_CLS(X), HFX_CLS(Y), GradReal>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, HFX_ROW_LDS_MAX);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((unsigned)nblocks, (unsigned)gnsplit(HFX_CLS(X), HFX_CLS(Y), sizeof(GradReal) == 4));
    hipLaunchKernelGGL((grad_row_kernel<HFX_CLS(X), HFX_CLS(Y), GradReal>), grid,
                       dim3(hfx_grad_threads(X, Y, sizeof(GradReal) == 4)), lds, s, a);
    return hipGetLastError();
}
typedef hipError_t (*grad_rows_fn)(const GradRowArgs&, int, hipStream_t);
const grad_rows_fn GRAD_ROWS_TABLE[6][6] = HFX_TABLE(grad_rows_xy);
#endif

#if HFX_PART == 6
template <int X, int Y>
hipError_t block_xy(const BlockArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((block_kernel<HFX_CLS(X), HFX_CLS(Y)>), dim3((unsigned)a.nq), dim3(64), 0, s, a);
    return hipGetLastError();
}
template <int X, int Y>
void row_shape_xy(int kind, int* nrow, int* nab) {
    constexpr int NA = ncart(pair_class_la(X)), NB = ncart(pair_class_lb(X));
    constexpr int NAt64 = split_a(HFX_CLS(X), HFX_CLS(Y), false) ? 1 : NA;
    constexpr int NAt32 = split_a(HFX_CLS(X), HFX_CLS(Y), true) ? 1 : NA;
    constexpr int NAtg64 = gsplit_a(HFX_CLS(X), HFX_CLS(Y), false) ? 1 : NA;
    constexpr int NAtg32 = gsplit_a(HFX_CLS(X), HFX_CLS(Y), true) ? 1 : NA;
    constexpr int NAtn64 = split_a_rows(HFX_CLS(X), HFX_CLS(Y), false, true) ? 1 : NA;
    constexpr int NAtn32 = split_a_rows(HFX_CLS(X), HFX_CLS(Y), true, true) ? 1 : NA;
    const int NAt = kind == 0 ? NAt64 : kind == 1 ? NAt32 : kind == 2 ? NAtg64 : kind == 3 ? NAtg32
                  : kind == 4 ? NAtn64 : NAtn32;
    *nrow = NAt + NB;
    *nab = NAt * NB;
}
typedef hipError_t (*blk_fn)(const BlockArgs&, hipStream_t);
typedef void (*shape_fn)(int, int*, int*);
const blk_fn BLK_TABLE[6][6] = HFX_TABLE(block_xy);
const shape_fn SHAPE_TABLE[6][6] = HFX_TABLE(row_shape_xy);
#endif

}  // namespace

#define HFX_CHECK_COMBO(cb, ck) \
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb) return hipErrorInvalidValue

#if HFX_PART == 1
hipError_t hfx_launch_jk(int cb, int ck, bool fp32, const JKArgs& a, int ntiles, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    if (ntiles <= 0) return hipErrorInvalidValue;
    return JK_TABLE[cb][ck](fp32, a, ntiles, s);
}
#endif

#if HFX_PART == 2
hipError_t hfx_launch_jk_rows_f64(int cb, int ck, bool narrow, const RowArgs& a, int nblocks, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    return (narrow ? ROWS_NARROW_TABLE : ROWS_TABLE)[cb][ck](a, nblocks, s);
}
#endif
#if HFX_PART == 3
hipError_t hfx_launch_jk_rows_f32(int cb, int ck, bool narrow, const RowArgs& a, int nblocks, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    return (narrow ? ROWS_NARROW_TABLE : ROWS_TABLE)[cb][ck](a, nblocks, s);
}
#endif
#if HFX_PART == 4
hipError_t hfx_launch_grad_f64(int cb, int ck, const GradArgs& a, int nblocks, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    return GRAD_TABLE[cb][ck](a, nblocks, s);
}
hipError_t hfx_launch_grad_rows_f64(int cb, int ck, const GradRowArgs& a, int nblocks, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    return GRAD_ROWS_TABLE[cb][ck](a, nblocks, s);
}
#endif
#if HFX_PART == 5
hipError_t hfx_launch_grad_f32(int cb, int ck, const GradArgs& a, int nblocks, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    return GRAD_TABLE[cb][ck](a, nblocks, s);
}
hipError_t hfx_launch_grad_rows_f32(int cb, int ck, const GradRowArgs& a, int nblocks, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    return GRAD_ROWS_TABLE[cb][ck](a, nblocks, s);
}
#endif

#if HFX_PART == 6
hipError_t hfx_launch_jk_rows_f64(int cb, int ck, bool narrow, const RowArgs& a, int nblocks, hipStream_t s);
hipError_t hfx_launch_jk_rows_f32(int cb, int ck, bool narrow, const RowArgs& a, int nblocks, hipStream_t s);
hipError_t hfx_launch_grad_f64(int cb, int ck, const GradArgs& a, int nblocks, hipStream_t s);
hipError_t hfx_launch_grad_f32(int cb, int ck, const GradArgs& a, int nblocks, hipStream_t s);
hipError_t hfx_launch_grad_rows_f64(int cb, int ck, const GradRowArgs& a, int nblocks, hipStream_t s);
hipError_t hfx_launch_grad_rows_f32(int cb, int ck, const GradRowArgs& a, int nblocks, hipStream_t s);

int hfx_nsplit(int cb, int ck) {
    return nsplit(pair_class_la(cb), pair_class_lb(cb), pair_class_la(ck), pair_class_lb(ck));
}

void hfx_row_shape(int cb, int ck, int kind, int* nrow, int* nab) {
    *nrow = *nab = 0;
    if (cb < 0 || cb > 5 || ck < 0 || ck > cb || kind < 0 || kind > 5) return;
    SHAPE_TABLE[cb][ck](kind, nrow, nab);
}

hipError_t hfx_launch_jk_rows(int cb, int ck, bool fp32, bool narrow, const RowArgs& a, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipErrorInvalidValue;
    return fp32 ? hfx_launch_jk_rows_f32(cb, ck, narrow, a, nblocks, s)
                : hfx_launch_jk_rows_f64(cb, ck, narrow, a, nblocks, s);
}

hipError_t hfx_launch_grad(int cb, int ck, bool fp32, const GradArgs& a, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipErrorInvalidValue;
    return fp32 ? hfx_launch_grad_f32(cb, ck, a, nblocks, s) : hfx_launch_grad_f64(cb, ck, a, nblocks, s);
}

hipError_t hfx_launch_grad_rows(int cb, int ck, bool fp32, const GradRowArgs& a, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipErrorInvalidValue;
    return fp32 ? hfx_launch_grad_rows_f32(cb, ck, a, nblocks, s) : hfx_launch_grad_rows_f64(cb, ck, a, nblocks, s);
}

hipError_t hfx_launch_block(int cb, int ck, const BlockArgs& a, hipStream_t s) {
    HFX_CHECK_COMBO(cb, ck);
    if (a.nq <= 0) return hipErrorInvalidValue;
    return BLK_TABLE[cb][ck](a, s);
}
hipError_t hfx_launch_schwarz(int cls, const ShellPair* pairs, const PrimPair* prims, const int* list, int npairs,
                              const double* scale, unsigned long long* q2bits, hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((npairs + 255) / 256);
    switch (cls) {
        case 0:
            hipLaunchKernelGGL((schwarz_kernel<0, 0>), dim3(nb, hfx_nsplit(0, 0)), dim3(256), 0, s, pairs, prims,
                               list, npairs, scale, q2bits);
            break;
        case 1:
            hipLaunchKernelGGL((schwarz_kernel<1, 0>), dim3(nb, hfx_nsplit(1, 1)), dim3(256), 0, s, pairs, prims,
                               list, npairs, scale, q2bits);
            break;
        case 2:
            hipLaunchKernelGGL((schwarz_kernel<1, 1>), dim3(nb, hfx_nsplit(2, 2)), dim3(256), 0, s, pairs, prims,
                               list, npairs, scale, q2bits);
            break;
        case 3:
            hipLaunchKernelGGL((schwarz_kernel<2, 0>), dim3(nb, hfx_nsplit(3, 3)), dim3(256), 0, s, pairs, prims,
                               list, npairs, scale, q2bits);
            break;
        case 4:
            hipLaunchKernelGGL((schwarz_kernel<2, 1>), dim3(nb, hfx_nsplit(4, 4)), dim3(256), 0, s, pairs, prims,
                               list, npairs, scale, q2bits);
            break;
        case 5:
            hipLaunchKernelGGL((schwarz_kernel<2, 2>), dim3(nb, hfx_nsplit(5, 5)), dim3(256), 0, s, pairs, prims,
                               list, npairs, scale, q2bits);
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

hipError_t hfx_launch_ket_gather(const KetTab& kt, int ncomp, int nd, const double* D, int n, hipStream_t s) {
    const size_t tot = (size_t)ncomp * kt.nk;
    if (tot == 0) return hipSuccess;
    unsigned nb = (unsigned)((tot + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(ket_gather_kernel, dim3(nb), dim3(256), 0, s, kt, ncomp, nd, D, n);
    return hipGetLastError();
}

hipError_t hfx_launch_ket_scatter(const KetTab& kt, int ncomp, int nd, double* Jb, int n, hipStream_t s) {
    const size_t tot = (size_t)ncomp * kt.nk;
    if (tot == 0) return hipSuccess;
    unsigned nb = (unsigned)((tot + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(ket_scatter_kernel, dim3(nb), dim3(256), 0, s, kt, ncomp, nd, Jb, n);
    return hipGetLastError();
}

// 0-based 8-fold packed index of (i j|k l) (the index of cints.c:327-347, 64-bit).
__device__ HFX_INLINE long long packed_ijkl(long long i, long long j, long long k, long long l) {
    const long long ij = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
    const long long kl = k >= l ? k * (k + 1) / 2 + l : l * (l + 1) / 2 + k;
    return ij >= kl ? ij * (ij + 1) / 2 + kl : kl * (kl + 1) / 2 + ij;
}

// X[((a n + b) n + c) n + d] = (a b|c d) from the packed vector.
__global__ __launch_bounds__(256) void unpack_int2_kernel(const double* __restrict__ int2, double* __restrict__ X,
                                                          int n) {
    const size_t n1 = n, tot = n1 * n1 * n1 * n1;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const size_t d = k % n1, c = (k / n1) % n1, b = (k / (n1 * n1)) % n1, a = k / (n1 * n1 * n1);
        X[k] = int2[packed_ijkl(a, b, c, d)];
    }
}

// out[packed(i j|k l)] = Y[((i n + j) n + k) n + l] for the canonical i >= j, k >= l, ij >= kl.
__global__ __launch_bounds__(256) void pack_int2_kernel(const double* __restrict__ Y, double* __restrict__ out, int n) {
    const size_t n1 = n, tot = n1 * n1 * n1 * n1;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const long long l = k % n1, kk = (k / n1) % n1, j = (k / (n1 * n1)) % n1, i = k / (n1 * n1 * n1);
        if (i < j || kk < l) continue;
        const long long ij = i * (i + 1) / 2 + j, kl = kk * (kk + 1) / 2 + l;
        if (ij < kl) continue;
        out[ij * (ij + 1) / 2 + kl] = Y[k];
    }
}

hipError_t hfx_launch_unpack_int2(const double* int2, double* X, int n, hipStream_t s) {
    hipLaunchKernelGGL(unpack_int2_kernel, dim3(4096), dim3(256), 0, s, int2, X, n);
    return hipGetLastError();
}

hipError_t hfx_launch_pack_int2(const double* Y, double* out, int n, hipStream_t s) {
    hipLaunchKernelGGL(pack_int2_kernel, dim3(4096), dim3(256), 0, s, Y, out, n);
    return hipGetLastError();
}

// One thread per shell pair (i, j): the largest |D| of the block, rounded up to
// float so that the screening test never drops a quartet it should keep.
__global__ __launch_bounds__(256) void shell_dmax_kernel(const double* __restrict__ D, int n,
                                                         const int* __restrict__ f0, const int* __restrict__ nc,
                                                         int nsh, float* __restrict__ dsh) {
    const size_t tot = (size_t)nsh * nsh;
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < tot; k += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(k / nsh), j = (int)(k - (size_t)i * nsh);
        double m = 0.0;
        for (int a = 0; a < nc[i]; a++)
            for (int b = 0; b < nc[j]; b++) m = fmax(m, fabs(D[(size_t)(f0[i] + a) * n + f0[j] + b]));
        dsh[k] = (float)(m * (1.0 + 1e-6));
    }
}

hipError_t hfx_launch_shell_dmax(const double* D, int n, const int* sh_f0, const int* sh_nc, int nsh, float* dsh,
                                 hipStream_t s) {
    const size_t tot = (size_t)nsh * nsh;
    unsigned nb = (unsigned)((tot + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(shell_dmax_kernel, dim3(nb), dim3(256), 0, s, D, n, sh_f0, sh_nc, nsh, dsh);
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

// Scratch of the allocation-free entry point (hfx_eri_shells_noalloc): the
// quartet plans and integral blocks of one call, in a file-scope device array
// of HFX_SCRATCH_BYTES (one copy per GPU, held until the process ends). In
// host code the array's name is a host-side stand-in, never a device address:
// hfx_scratch_address looks the address up once per device, with that device
// current, and callers use only the address it returns.
static __device__ __align__(256) unsigned char scratch[HFX_SCRATCH_BYTES];

hipError_t hfx_scratch_address(int device, void** ptr) {
    static std::mutex mu;
    static void* addr[HFX_MAX_DEVICES] = {};
    if (!ptr || device < 0 || device >= HFX_MAX_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

#endif  // HFX_PART == 6

#if HFX_PART == 7
namespace {

// One wavefront per shell pair (grid-stride over the pair list). Overlap and
// kinetic terms: lanes over the primitive pairs; nuclear attraction: lanes over
// the nuclei, primitive pairs in a wavefront-uniform loop.
template <int LA, int LB, bool DERIV>
__global__ __launch_bounds__(256) void one_e_kernel(OneEArgs a) {
    constexpr int NA = ncart(LA), NB = ncart(LB);
    extern __shared__ double gl[];   // 3 nat (gradient)
    if constexpr (DERIV) {
        for (int k = threadIdx.x; k < 3 * a.nat; k += blockDim.x) gl[k] = 0.0;
        __syncthreads();
    }
    const int lane = threadIdx.x % 64;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int nwave = gridDim.x * blockDim.x / 64;
    const RysTables tab = dev_tables();
    for (int pi = wave; pi < a.npairs; pi += nwave) {
        const int A = a.pairA[pi], B = a.pairB[pi];
        const int ata = a.sh_atom[A], atb = a.sh_atom[B];
        const double* ra = a.xyz + 3 * ata;
        const double* rb = a.xyz + 3 * atb;
        const double AB[3] = {ra[0] - rb[0], ra[1] - rb[1], ra[2] - rb[2]};
        const double r2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        const int na = a.sh_nprim[A], nb = a.sh_nprim[B], pa0 = a.sh_prim0[A], pb0 = a.sh_prim0[B];
        const int fa = a.sh_f0[A], fb = a.sh_f0[B];
        const double poly = pow(1.0 + sqrt(r2), LA + LB + 1);
        // primitive pair k -> geometry; false when all its terms are below the screen
        auto prim = [&](int k, OePrim& q) -> bool {
            const int ia = k / nb, ib = k - ia * nb;
            q.a = a.alpha[pa0 + ia];
            q.b = a.alpha[pb0 + ib];
            q.p = q.a + q.b;
            q.oo2p = 0.5 / q.p;
            const double mu_r2 = q.a * q.b / q.p * r2;
            if (mu_r2 > 745.0) return false;
            q.K = a.coef[pa0 + ia] * a.coef[pb0 + ib] * exp(-mu_r2);
            const double bound = fabs(q.K) * fmax(HFX_PI / q.p * sqrt(HFX_PI / q.p), 2.0 * HFX_PI / q.p * a.ztot) *
                                 (1.0 + q.p) * (1.0 + q.p) * poly;
            if (bound < HFX_ONE_E_SCREEN) return false;
            const double ip = 1.0 / q.p;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                q.P[d] = (q.a * ra[d] + q.b * rb[d]) * ip;
                q.PA[d] = q.P[d] - ra[d];
                q.PB[d] = q.P[d] - rb[d];
            }
            return true;
        };
        if constexpr (DERIV) {
            double pw[NA][NB], ww[NA][NB];
#pragma unroll
            for (int i = 0; i < NA; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const size_t ij = (size_t)(fa + i) + (size_t)(fb + j) * a.n;
                    const double sc = a.scale[fa + i] * a.scale[fb + j];
                    pw[i][j] = a.P[ij] * sc;
                    ww[i][j] = a.W[ij] * sc;
                }
            double gA[3] = {0.0, 0.0, 0.0};
            for (int k = lane; k < na * nb; k += 64) {
                OePrim q;
                if (prim(k, q)) oe_st_grad<LA, LB>(q, pw, ww, gA);
            }
            for (int c0 = 0; c0 < a.nat; c0 += 64) {
                const int c = c0 + lane;
                const bool on = c < a.nat;
                const double Cc[3] = {on ? a.xyz[3 * c] : 0.0, on ? a.xyz[3 * c + 1] : 0.0,
                                      on ? a.xyz[3 * c + 2] : 0.0};
                const double Zc = on ? a.Z[c] : 0.0;
                double h[3] = {0.0, 0.0, 0.0};
                for (int k = 0; k < na * nb; k++) {
                    OePrim q;
                    if (!prim(k, q)) continue;   // wavefront-uniform
                    if (on) oe_v_grad<LA, LB>(q, Cc, Zc, AB, tab, pw, h);
                }
                if (on) {
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        // derivative on the centre of A (the other half is the (B, A)
                        // visit's) and the Hellmann-Feynman term of nucleus c
                        gA[d] += h[d];
                        if (h[d] != 0.0) atomicAdd(&gl[3 * c + d], -2.0 * h[d]);
                    }
                }
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double v = wave_sum(2.0 * gA[d]);
                if (lane == 0 && v != 0.0) atomicAdd(&gl[3 * ata + d], v);
            }
        } else {
            double S[NA][NB], T[NA][NB];
#pragma unroll
            for (int i = 0; i < NA; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) S[i][j] = T[i][j] = 0.0;
            for (int k = lane; k < na * nb; k += 64) {
                OePrim q;
                if (prim(k, q)) oe_st<LA, LB>(q, S, T);
            }
#pragma unroll
            for (int i = 0; i < NA; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const double s_ = wave_sum(S[i][j]), t_ = wave_sum(T[i][j]);
                    if (lane == 0) {
                        const int fi = fa + i, fj = fb + j;
                        const double sc = a.scale[fi] * a.scale[fj];
                        const size_t ij = (size_t)fi + (size_t)fj * a.n, ji = (size_t)fj + (size_t)fi * a.n;
                        if (a.S) a.S[ij] = a.S[ji] = sc * s_;
                        if (a.T) a.T[ij] = a.T[ji] = sc * t_;
                    }
                }
            double V[NA][NB];
#pragma unroll
            for (int i = 0; i < NA; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) V[i][j] = 0.0;
            for (int c0 = 0; c0 < a.nat; c0 += 64) {
                const int c = c0 + lane;
                const bool on = c < a.nat;
                const double Cc[3] = {on ? a.xyz[3 * c] : 0.0, on ? a.xyz[3 * c + 1] : 0.0,
                                      on ? a.xyz[3 * c + 2] : 0.0};
                const double Zc = on ? a.Z[c] : 0.0;
                for (int k = 0; k < na * nb; k++) {
                    OePrim q;
                    if (!prim(k, q)) continue;   // wavefront-uniform
                    if (on) oe_v<LA, LB>(q, Cc, Zc, AB, tab, V);
                }
            }
#pragma unroll
            for (int i = 0; i < NA; i++)
#pragma unroll
                for (int j = 0; j < NB; j++) {
                    const double v_ = wave_sum(V[i][j]);
                    if (lane == 0 && a.V) {
                        const int fi = fa + i, fj = fb + j;
                        const double sc = a.scale[fi] * a.scale[fj];
                        a.V[(size_t)fi + (size_t)fj * a.n] = a.V[(size_t)fj + (size_t)fi * a.n] = sc * v_;
                    }
                }
        }
    }
    if constexpr (DERIV) {
        __syncthreads();
        for (int k = threadIdx.x; k < 3 * a.nat; k += blockDim.x)
            if (gl[k] != 0.0) unsafeAtomicAdd(&a.grad[k], gl[k]);
    }
}

template <int LA, int LB>
hipError_t one_e_l(bool deriv, const OneEArgs& a, hipStream_t s) {
    if (a.npairs <= 0) return hipSuccess;
    unsigned nb = (unsigned)((a.npairs + 3) / 4);   // four wavefronts per block
    if (nb > 2048) nb = 2048;
    if (deriv) {
        const size_t lds = 3 * (size_t)a.nat * sizeof(double);
        if (lds > HFX_ROW_LDS_MAX) return hipErrorInvalidValue;
        static bool attr_set = false;
        if (!attr_set) {
            hipError_t e = hipFuncSetAttribute((const void*)one_e_kernel<LA, LB, true>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, HFX_ROW_LDS_MAX);
            if (e != hipSuccess) return e;
            attr_set = true;
        }
        hipLaunchKernelGGL((one_e_kernel<LA, LB, true>), dim3(nb), dim3(256), lds, s, a);
    } else {
        hipLaunchKernelGGL((one_e_kernel<LA, LB, false>), dim3(nb), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

typedef hipError_t (*one_e_fn)(bool, const OneEArgs&, hipStream_t);
const one_e_fn ONE_E_TABLE[3][3] = {{one_e_l<0, 0>, one_e_l<0, 1>, one_e_l<0, 2>},
                                    {one_e_l<1, 0>, one_e_l<1, 1>, one_e_l<1, 2>},
                                    {one_e_l<2, 0>, one_e_l<2, 1>, one_e_l<2, 2>}};

}  // namespace

hipError_t hfx_launch_one_e(int la, int lb, bool deriv, const OneEArgs& a, hipStream_t s) {
    if (la < 0 || la > 2 || lb < 0 || lb > 2) return hipErrorInvalidValue;
    return ONE_E_TABLE[la][lb](deriv, a, s);
}
#endif  // HFX_PART == 7
