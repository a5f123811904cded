// Interface between the host runtime (hfx_host.cpp) and the HIP kernels
// (hfx_kernels.hip). Not part of the public ABI.
#pragma once

#include <array>
#include <vector>

#include <hip/hip_runtime.h>

#include "hfx_common.h"
#include "rys_params.h"

// Arguments of one J/K launch: bra pairs of class X (rows), ket pairs of class
// Y (columns), X >= Y; both lists sorted by Schwarz bound, descending, so the
// significant kets of a row form a prefix. Row i evaluates columns
// [kbeg[i], kend[i]) (one precision's share of that prefix; for X == Y the
// prefix is cut at the diagonal). The ket list is one of the class's
// sublists of equal primitive-pair count, so the primitive loops of a
// wavefront have one trip count. Tiles: 4 rows x 64 columns, 256 threads.
struct JKArgs {
    const ShellPair* pairs;
    const PrimPair* prims;
    const int* bra_list;
    const int* ket_list;
    const int* tile_i0;
    const int* tile_j0;
    const int* kbeg;
    const int* kend;
    int nbra;
    const double* D;  // scaled, symmetrised density, n x n
    double* J;        // unsymmetrised accumulators, n x n
    double* K;
    int n;
};

static constexpr int HFX_SUB_MAX = 16;  // primitive-count sublists of a ket class
static constexpr int HFX_WIN_MAX = 16;  // column windows of a ket class

// Per-position record of a ket table: pair id, first function of C and of D,
// and cp = cpa | cpb << 15 | (C == D) << 30 (cpa / cpb: LDS column of the
// first function of C / of D among the functions of the L_C / L_D shells).
struct alignas(16) KetInfo {
    int pid, fa, fb, cp;
};

// Ket tables of one pair class: all its significant pairs, grouped into
// sublists by primitive-pair count rounded up to a bucket size np[s] (position
// k in [off[s], off[s+1]) belongs to sublist s, Q-descending within it; a ket
// with fewer primitive pairs is padded with copies of its last one with K = 0),
// stored field-major so that consecutive positions (consecutive lanes) read
// consecutive addresses.
struct KetTab {
    const KetInfo* info;
    const double* cd;     // [3][nk]: C - D
    const double* prim;   // sublist s at prim + psoff[s]: [np_s][10 planes][pitch] (KetChunk)
    const float* primf;   // the same in single precision (FP32 kernels), sublist s at primf + psoff_f[s]
    double* Dcd;          // [NC * ND][nk]: density blocks D_cd, gathered per build
    double* Jcd;          // [NC * ND][nk]: J_cd accumulators, scattered into J after the build
    const int* shcd;      // [nk]: shells C | D << 16 (density-weighted screening)
    const float* qk;      // [nk]: Schwarz bound of the ket pair (density-weighted screening)
    int nk;
    int S;
    int off[HFX_SUB_MAX + 1];
    int np[HFX_SUB_MAX];
    long long psoff[HFX_SUB_MAX];
    long long psoff_f[HFX_SUB_MAX];
};

// ---- quartet lists built on the GPU (hfx_lists.hip) ----
// Significant pairs in class order (Q-descending within a class) and in
// (class, column window, primitive-count bucket) order; runs = (class, window,
// bucket, count) of the second order; pos = position of every significant pair
// in its class list.
struct PairOrder {
    int* cls_order = nullptr;   // [npairs] (device, allocated by the caller)
    int* sub_order = nullptr;   // [npairs]
    int* pos = nullptr;         // [npairs], -1 for pairs that are not significant
    int ncls[8] = {0};
    int nsig = 0;
    std::vector<std::array<int, 4>> runs;
};
// Mutable views of one class's ket table for the fill kernel.
struct KetOut {
    KetInfo* info;
    double* cd;
    int* shcd;
    float* qk;
    double* prim;
    float* primf;
};
// Per ket of one class table: bound, position in the class list, primitive
// count ([nk + 1], the last 0), and the exclusive prefix sums of the counts
// over the table ([nk + 1]).
struct KetAux {
    double* q;
    int* pos;
    int* np;
    int* cnp;
};
// Ket ranges of one class combination (bra rows = the bra class list).
struct RangeArgs {
    const ShellPair* pairs;
    const int* bra_list;
    KetTab kt;
    KetAux x;
    int nb, S;
    int same;            // bra class == ket class: kets up to the diagonal
    double thr, thr32;
    int* rkb[2];         // [nb * S] per precision
    int* rke[2];
    double* stats;       // [S][nq64, nprimq64, nq32, nprimq32]
};
// Row-mode blocks of one class combination and precision.
struct BlockArgsL {
    const int* rkb;
    const int* rke;
    int nb, S, R, chunk;
};
// d_pwin: column window of every pair (nullptr: one window)
hipError_t hfx_pair_order(const ShellPair* d_pairs, const int* d_pwin, int np, double thr, PairOrder& o,
                          hipStream_t s);
hipError_t hfx_launch_ket_fill(const KetTab& t, const KetOut& o, const KetAux& x, const int* sub_order,
                               const ShellPair* pairs, const PrimPair* prims, const PrimExp* pexp, const int* pos,
                               hipStream_t s);
hipError_t hfx_launch_row_ranges(const RangeArgs& a, hipStream_t s);
hipError_t hfx_launch_group_spans(const BlockArgsL& a, double* sums, hipStream_t s);
hipError_t hfx_group_blocks(const BlockArgsL& a, int** brow, int** bk0, int* nblk, std::vector<void*>& allocs,
                            hipStream_t s);

// Arguments of one row-mode J/K launch (bra class X, ket class Y, X >= Y):
// block b owns bra rows blk_row[b] .. blk_row[b] + R - 1 and the ket
// positions [blk_k0[b], blk_k0[b] + chunk), intersected, sublist by sublist,
// with each row's range [kbeg[i*S + s], kend[i*S + s]) (one precision's share
// of the row's Schwarz prefix; for X == Y cut at the diagonal). The K
// contributions of each row land in the rows of its shells A and B,
// accumulated in LDS over all the block's kets and flushed once: LDS columns
// are the functions of the L_C shells (then, when L_D != L_C, those of the
// L_D shells from column offd on), or with column windows the window's range
// of them (from col0_c / col0_d); bf_c / bf_d map columns back to basis
// functions. With dlds the same rows of the density are staged in LDS in the
// same layout. J_ab is reduced over the block; J_cd of each ket is summed over
// the R rows and added to the ket table's accumulators.
struct RowArgs {
    const ShellPair* pairs;
    const PrimPair* prims;
    const PrimPairF* primsf;   // single-precision copy of prims (FP32 launches)
    const int* bra_list;
    KetTab kt;
    const int* kbeg;
    const int* kend;
    const int* blk_row;   // first of the block's R consecutive bra rows
    const int* blk_k0;
    int chunk;
    int R;                // bra rows per block (1..HFX_ROWS_MAX)
    int nbra;
    int ncols;
    int offd;
    const int* bf_c;
    const int* bf_d;
    // column window of the ket table: LDS column 0 (offd) is position col0_c
    // (col0_d) among the L_C (L_D) functions; both 0 without windows
    int col0_c, col0_d;
    int dlds;       // stage the density rows of the bra shells in LDS
    const double* D;
    double* J;
    double* K;
    int n;
    // density-weighted screening (dsh == nullptr: off): a quartet is skipped when
    // Q_ab Q_cd max(dsh over its six shell-pair blocks) < dthr, dsh[i * nsh + j]
    // = max |D| over the block of shells i, j
    const float* dsh;
    float dthr;
    int nsh;
};

// Bytes of dynamic LDS of a row-mode launch: K rows [R][nrow][ncols] and J_ab
// [R][nab] in double, the optional density rows [R][nrow][ncols] in the
// arithmetic type.
inline size_t hfx_row_lds(int nrow, int nab, int ncols, int R, int dlds, bool fp32) {
    size_t b = (size_t)R * ((size_t)nrow * ncols + nab) * sizeof(double);
    if (dlds) {
        const size_t m = (size_t)R * nrow * ncols;
        b += fp32 ? ((m + 1) / 2) * sizeof(double) : m * sizeof(double);
    }
    return b;
}

// Arguments of one gradient launch (same row blocks as RowArgs): every quartet's
// sum_abcd gam_abcd d(ab|cd)/dR goes to the atoms of its four shells, summed in
// LDS per block and flushed into grad[3 * nat].
struct GradArgs {
    const ShellPair* pairs;
    const PrimPair* prims;
    const PrimExp* pexp;
    const int* bra_list;
    const int* ket_list;
    const int* kbeg;
    const int* kend;
    const int* blk_row;
    const int* blk_k0;
    int chunk;
    const int* shell_atom;
    int nat;
    const double* D;
    double* grad;
    int n;
};

// Arguments of one row-mode gradient launch: the blocks, rows, ket table and
// LDS density rows of the row-mode J/K launch of the same class combination
// (RowArgs); the nine derivative components of every quartet go to the atoms
// of its shells (bf_atom of their first functions), summed in LDS per block
// (3 nat doubles after the density rows) and flushed into grad[3 nat].
struct GradRowArgs {
    const ShellPair* pairs;
    const PrimPair* prims;
    const PrimExp* pexp;
    const int* bra_list;
    KetTab kt;
    const int* kbeg;
    const int* kend;
    const int* blk_row;
    const int* blk_k0;
    int chunk;
    int R;
    int nbra;
    int ncols;
    int offd;
    const int* bf_c;
    const int* bf_d;
    // column window of the ket table: LDS column 0 (offd) is position col0_c
    // (col0_d) among the L_C (L_D) functions; both 0 without windows
    int col0_c, col0_d;
    int dlds;
    const int* bf_atom;
    int nat;
    const double* D;
    double* grad;
    int n;
};

// Arguments of the one-electron kernels: a list of shell pairs (A, B) with
// l_A = LA, l_B = LB, one wavefront per pair. Without derivatives the S, T, V
// blocks (and their transposes) are written into column-major n x n matrices;
// with derivatives the gradient of sum P (T + V) - sum W S with respect to the
// centre of A (over all ordered pairs) and the Hellmann-Feynman term of every
// nucleus are summed into grad[3 nat] (3 nat doubles of LDS per block).
struct OneEArgs {
    const int* pairA;
    const int* pairB;
    int npairs;
    const int* sh_atom;
    const int* sh_f0;
    const int* sh_nprim;
    const int* sh_prim0;
    const double* alpha;   // shell-level primitives
    const double* coef;
    const double* xyz;     // 3 nat, bohr
    const double* Z;       // nat
    int nat;
    double ztot;           // sum of |Z| (screening bound)
    const double* scale;   // per function
    int n;
    double* S;             // integrals
    double* T;
    double* V;
    const double* P;       // gradient
    const double* W;
    double* grad;
};
static constexpr double HFX_ONE_E_SCREEN = 1e-17;   // primitive-pair bound below which terms are dropped

// Arguments of the explicit-quartet kernel: quartet q uses pairs bra[q], ket[q]
// (already oriented l_A >= l_B, l_C >= l_D, class(bra) >= class(ket)) and
// writes its block at out + off[q] in the (a, b, c, d) order of the oriented
// pairs.
struct BlockArgs {
    const ShellPair* pairs;
    const PrimPair* prims;
    const int* bra;
    const int* ket;
    const long long* off;
    const int* qidx;   // quartet ids of this class combination
    int nq;
    double* out;
    // packed mode (off unused): every integral, times the scales of its four
    // functions, goes to out[index(i, j, k, l)] of the 8-fold packed array
    int packed;
    const double* scale;
};

static constexpr int HFX_ROWS_MAX = 4;   // bra rows per row-mode block
// Threads of a J/K row-mode block of class (X|Y) (pair classes) in FP32 or FP64:
// 256, or 512 / 1024 for the classes whose K and density rows hold a CU to one
// or two blocks while their registers leave room for two or four wavefronts per
// SIMD (chosen per class from launch times at 64 waters 6-31G**, n = 1600).
HFX_HD constexpr int hfx_row_threads(int X, int Y, bool fp32) {
    if (fp32) {
        if (X == 3 && Y == 1) return 1024;                            // (20|10)
        if ((X == 1 && Y == 1) || (X == 2 && Y == 1) || (X == 4 && Y == 1) || (X == 4 && Y == 3) ||
            (X == 5 && Y == 4) || (X == 5 && Y == 0))
            return 512;                                    // (10|10) (11|10) (21|10) (21|20) (22|21) (22|00)
        return 256;
    }
    if ((X == 2 && Y == 1) || (X == 3 && Y == 3) || (X == 5 && Y == 3) || (X == 3 && Y == 2))
        return 512;                                        // (11|10) (20|20) (22|20) (20|11)
    return 256;
}
// Whether a row-mode J/K launch of class (X|Y) stages the density rows of its
// bra shells in LDS (when they fit the LDS budget). Not for these classes,
// whose launches measured faster reading the density through the caches
// (64 waters, launches serialised: FP64 (10|10) -1 % in the FP64 build and
// -16 % in the mixed build, (11|11) -4 %, (21|21) -10 %, (22|11) -1 %,
// (22|22) -9 %, (20|10) -17 % (its rows take 136 KB with the density, one
// block per CU); FP32 (22|22) -12 %); every other class is faster staging them.
HFX_HD constexpr bool hfx_row_stage_density(int X, int Y, bool fp32) {
    if (fp32) return !(X == 5 && Y == 5);
    return !((X == 1 && Y == 1) || (X == 2 && Y == 2) || (X == 4 && Y == 4) || (X == 5 && Y == 2) ||
             (X == 5 && Y == 5) || (X == 3 && Y == 1));
}
// Threads of a gradient row-mode block of class (X|Y) in FP32 or FP64: 256,
// or 512 for the FP32 classes (21|11), (21|10), (20|11) (10-24 % faster per
// class at 64 waters; every other class is slower with 512 or 1024).
HFX_HD constexpr int hfx_grad_threads(int X, int Y, bool fp32) {
    return fp32 && ((X == 4 && Y == 2) || (X == 4 && Y == 1) || (X == 3 && Y == 2)) ? 512 : 256;
}
static constexpr int HFX_AUX_STREAMS = 3; // extra streams of the class launches (at most)
// dynamic LDS of a row-mode block: gfx950 gives a workgroup up to 160 KB, of
// which 4 KB stay for the kernels' static arrays
static constexpr size_t HFX_ROW_LDS_MAX = 156 * 1024;
static constexpr int HFX_TILE_ROWS = 4;   // one bra row per wavefront
static constexpr int HFX_TILE_COLS = 64;

hipError_t hfx_launch_jk(int cls_bra, int cls_ket, bool fp32, const JKArgs& a, int ntiles, hipStream_t s);
// LDS rows per bra row (nrow) and J_ab entries (nab) of a row-mode launch:
// kind 0 = J/K in FP64, 1 = J/K in FP32, 2 = gradient in FP64, 3 = gradient
// in FP32, 4 / 5 = narrow J/K in FP64 / FP32 (their slicing differs)
void hfx_row_shape(int cls_bra, int cls_ket, int kind, int* nrow, int* nab);
// narrow: the launch sliced by split_a_rows / split_c_rows (eri_rys.h)
hipError_t hfx_launch_jk_rows(int cls_bra, int cls_ket, bool fp32, bool narrow, const RowArgs& a, int nblocks,
                              hipStream_t s);
hipError_t hfx_launch_grad(int cls_bra, int cls_ket, bool fp32, const GradArgs& a, int nblocks, hipStream_t s);
// LDS bytes of a row-mode gradient launch (density rows as in hfx_row_lds, then 3 nat doubles)
inline size_t hfx_grad_row_lds(int nrow, int ncols, int R, int dlds, bool fp32, int nat) {
    size_t b = 0;
    if (dlds) {
        const size_t m = (size_t)R * nrow * ncols;
        b += fp32 ? ((m + 1) / 2) * sizeof(double) : m * sizeof(double);
    }
    return b + 3 * (size_t)nat * sizeof(double);
}
hipError_t hfx_launch_grad_rows(int cls_bra, int cls_ket, bool fp32, const GradRowArgs& a, int nblocks,
                                hipStream_t s);
hipError_t hfx_launch_block(int cls_bra, int cls_ket, const BlockArgs& a, hipStream_t s);
// LDS limit of the fallback gradient kernel (3 nat doubles of per-atom sums)
static constexpr size_t HFX_GRAD_LDS_MAX = 128 * 1024;
// Device scratch of the allocation-free entry point: a file-scope array of
// HFX_SCRATCH_BYTES per GPU. hfx_scratch_address returns its device address
// for `device`, which must be the current device (looked up once per device).
static constexpr size_t HFX_SCRATCH_BYTES = (size_t)16 << 20;
static constexpr int HFX_MAX_DEVICES = 64;
// largest shell-quartet block, (dd|dd): 6^4 integrals
static constexpr int HFX_MAX_BLOCK = ((HFX_LMAX + 1) * (HFX_LMAX + 2) / 2) * ((HFX_LMAX + 1) * (HFX_LMAX + 2) / 2) *
                                     ((HFX_LMAX + 1) * (HFX_LMAX + 2) / 2) * ((HFX_LMAX + 1) * (HFX_LMAX + 2) / 2);
hipError_t hfx_scratch_address(int device, void** ptr);
// Q2[pair] = max over the diagonal integrals (ab|ab) of the pair (squared bound)
hipError_t hfx_launch_schwarz(int cls, const ShellPair* pairs, const PrimPair* prims, const int* list, int npairs,
                              const double* scale, unsigned long long* q2bits, hipStream_t s);
hipError_t hfx_launch_prep(const double* P, const double* scale, double* D, double* Jb, double* Kb, int n,
                           hipStream_t s);
// kt.Dcd <- D blocks of every ket of the table, kt.Jcd <- 0 (ncomp = NC * ND)
hipError_t hfx_launch_ket_gather(const KetTab& kt, int ncomp, int nd, const double* D, int n, hipStream_t s);
// Jb block of every ket += kt.Jcd
hipError_t hfx_launch_ket_scatter(const KetTab& kt, int ncomp, int nd, double* Jb, int n, hipStream_t s);
hipError_t hfx_launch_finish(const double* Jb, const double* Kb, const double* scale, double* G, double* J, double* K,
                             int n, hipStream_t s);
// X (n^4, [a][b][c][d]) <- packed 8-fold int2; packed <- canonical elements of Y
hipError_t hfx_launch_unpack_int2(const double* int2, double* X, int n, hipStream_t s);
hipError_t hfx_launch_pack_int2(const double* Y, double* out, int n, hipStream_t s);
// dsh[i * nsh + j] = max |D| over the block of shells i, j (rounded up to float)
hipError_t hfx_launch_shell_dmax(const double* D, int n, const int* sh_f0, const int* sh_nc, int nsh, float* dsh,
                                 hipStream_t s);
int hfx_nsplit(int cls_bra, int cls_ket);
hipError_t hfx_launch_one_e(int la, int lb, bool deriv, const OneEArgs& a, hipStream_t s);
