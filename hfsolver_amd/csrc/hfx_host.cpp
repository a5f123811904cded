// Host runtime of the direct J/K engine: the C ABI of include/hfx.h.
//
// Set-up (hfx_create / hfx_set_geometry):
//   1. regroup the per-basis-function arrays of gaussints (src/basis.f90:541-567)
//      into Cartesian shells, as src/gaussians_libint.f90:59-122 recovers shells
//      by stepping (L+1)(L+2)/2 functions; each function's contraction
//      coefficients are (shell coefficients) x (per-function scale);
//   2. all shell pairs (A,B), oriented l_A >= l_B, with their primitive pairs
//      (negligible primitive pairs screened out);
//   3. Schwarz bounds Q_AB = sqrt(max |(ab|ab)|) on the GPU;
//   4. on the GPU (hfx_lists.hip): per pair class, the significant pairs
//      sorted by Q (descending) and the class's chunked ket table (sublists by
//      primitive count); for every class combination each bra row's range of
//      kets per sublist with Q_ab Q_cd >= schwarz_threshold, split into an
//      FP64 and an FP32 share at fp32_threshold, the quartet counts, and the
//      row-mode block list; the host sizes the allocations from counts. Pair
//      classes with many ket columns get one ket table per column window (the
//      columns in the spatial order of their atoms), so that a row launch keeps
//      only its window's columns in LDS; a combination whose K rows still do
//      not fit runs as narrow row launches (sliced over the a-components); tile
//      lists are built only where even those do not fit.
// Per Fock build (hfx_fock_2e): P -> scaled symmetric D, D_cd gathered per ket,
// one launch per class combination and precision, J_cd scattered, then J, K
// symmetrised and unscaled, G = J - K/2. One-electron integrals and their
// gradient, the two-electron gradient and the store mode run on the GPU too.
#include <hip/hip_runtime.h>
#include <omp.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hfx.h"
#include "hfx_common.h"
#include "hfx_internal.h"
#include "rys.h"

namespace {

struct Shell {
    int L;
    int atom;
    int f0;       // first basis function
    int nprim;
    int prim0;    // into alpha / coef
    int cpos;     // first function among the functions of shells of the same L
};

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct hfx_engine {
    hfx_opts opts;
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    void* scratch = nullptr;   // device address of this GPU's allocation-free scratch array
    hfx_stats grad_stats{};    // statistics of the last hfx_grad_2e call
    // basis
    int nbf = 0;
    int nat = 0;
    std::vector<int> atno;
    std::vector<double> xyz;          // 3*nat
    double ztot = 0;                  // sum of the nuclear charges
    std::vector<Shell> shells;
    std::vector<double> alpha, coef;  // shell-level primitives
    std::vector<double> scale;        // per basis function
    std::vector<int> bf_atom;
    std::vector<int> bfL[HFX_LMAX + 1];   // functions of the shells of each L, in shell order
    int* d_bfL[HFX_LMAX + 1] = {nullptr};
    // row-mode LDS per block (a combination whose K rows need more runs as
    // narrow rows or, failing those, as tiles; the heaviest classes run one
    // block per CU)
    size_t lds_budget = HFX_ROW_LDS_MAX;
    // largest row-mode LDS per block with the density rows staged, FP64 and FP32
    // launches (both the whole row budget; the classes that read the density
    // through the caches instead: hfx_row_stage_density)
    size_t dlds_budget = HFX_ROW_LDS_MAX;
    size_t dlds_budget32 = HFX_ROW_LDS_MAX;
    // pairs
    std::vector<ShellPair> pairs;
    std::vector<PrimPair> prims;
    std::vector<PrimExp> pexp;
    std::vector<int> cls_list[HFX_NPAIRCLASS];  // significant pairs, Q descending
    struct Combo {
        int cb, ck, win = 0;
        int* d_ket = nullptr;   // ket sublist (equal primitive-pair count)
        int nket = 0;
        int npk = 0;            // its primitive-pair count
        int ntiles[2] = {0, 0};
        int* d_ti0[2] = {nullptr, nullptr};
        int* d_tj0[2] = {nullptr, nullptr};
        int* d_kbeg[2] = {nullptr, nullptr};
        int* d_kend[2] = {nullptr, nullptr};
        double nq[2] = {0, 0};
        double nprimq[2] = {0, 0};
        bool rowmode = false;   // J/K of (cb, ck) by the row-mode kernel (RowCombo)
        // one-row blocks of the gradient kernel
        int gnblk[2] = {0, 0};
        int gchunk[2] = {0, 0};
        int* d_gblk_row[2] = {nullptr, nullptr};
        int* d_gblk_k0[2] = {nullptr, nullptr};
    };
    std::vector<Combo> combos;
    // ket tables of every pair class and column window (sublists in ascending
    // primitive-pair count)
    struct KetTable {
        KetTab t{};
        int ncomp = 0, nd = 0;
    };
    KetTable ktab[HFX_NPAIRCLASS][HFX_WIN_MAX];
    // column windows (large molecules): window w of pair class c holds the
    // pairs whose C shell lies in the w-th of nwin[c] equal ranges of the L_C
    // columns, which then follow the atoms in space (Morton order); a row
    // launch over one window keeps only that window's columns in LDS. The
    // LDS columns of window w: [win_c0, win_c0 + win_cn) of the L_C functions,
    // then [win_d0, win_d0 + win_dn) of the L_D functions (L_D != L_C), or one
    // range [win_c0, win_c0 + win_cn) holding both (L_D == L_C).
    int win_cols = 1280;   // widest column range before a class is cut into windows (HFX_ROW_WINDOW_COLS, 0: never)
    int nwin[HFX_NPAIRCLASS] = {1, 1, 1, 1, 1, 1};
    int win_c0[HFX_NPAIRCLASS][HFX_WIN_MAX] = {}, win_cn[HFX_NPAIRCLASS][HFX_WIN_MAX] = {};
    int win_d0[HFX_NPAIRCLASS][HFX_WIN_MAX] = {}, win_dn[HFX_NPAIRCLASS][HFX_WIN_MAX] = {};
    // row-mode J/K of one class combination over the whole ket table
    struct RowCombo {
        int cb, ck, win = 0;
        int col0_c = 0, col0_d = 0;   // the window's first LDS columns (RowArgs)
        int R = 1, dlds = 0, ncols = 0, offd = 0;   // dlds: bit 0 FP64 launch, bit 1 FP32 launch
        bool narrow = false;   // launches sliced by split_a_rows / split_c_rows (K rows too wide otherwise)
        bool jk = true, grad = true;   // used by the J/K launches / by the derivative launches
        int nblk[2] = {0, 0};
        int chunk[2] = {0, 0};
        double cost[2] = {0, 0};   // modelled flops of the launch (orders the launches)
        int* d_kbeg[2] = {nullptr, nullptr};
        int* d_kend[2] = {nullptr, nullptr};
        int* d_blk_row[2] = {nullptr, nullptr};
        int* d_blk_k0[2] = {nullptr, nullptr};
    };
    std::vector<RowCombo> rcombos;
    // tuning of the row-mode blocks (HFX_ROW_RMAX, HFX_ROW_PAIR_KB, HFX_ROW_DLDS)
    int row_rmax = HFX_ROWS_MAX;
    size_t row_pair_bytes = 32 * 1024;
    double row_blocks = 128;   // blocks a launch is cut into when its rows are few (HFX_ROW_BLOCKS)
    rocblas_handle blas = nullptr;   // created on first use (AO -> MO transform)
    // density-weighted screening of the Fock builds (hfx_set_density_screening)
    double dthr = 0.0;
    float* d_dsh = nullptr;
    int* d_dsh_f0 = nullptr;
    int* d_dsh_nc = nullptr;
    int row_dlds = 1;
    int row_narrow = 1;   // narrow row launches where the K rows do not fit (HFX_ROW_NARROW)
    double npairs_sig = 0;
    // device memory
    ShellPair* d_pairs = nullptr;
    PrimPair* d_prims = nullptr;
    PrimPairF* d_primsf = nullptr;
    PrimExp* d_pexp = nullptr;
    int* d_cls_list[HFX_NPAIRCLASS] = {nullptr};
    double* d_scale = nullptr;
    double* d_D = nullptr;
    double* d_Jb = nullptr;
    double* d_Kb = nullptr;
    double* d_P = nullptr;    // host-pointer entry staging
    double* d_G = nullptr;
    double* d_J = nullptr;
    double* d_K = nullptr;
    int* d_shell_atom = nullptr;
    int* d_bf_atom = nullptr;
    // one-electron kernels: shell data, nuclei, shell-pair lists [ordered][l_A][l_B]
    struct OneEList {
        int* A = nullptr;
        int* B = nullptr;
        int n = 0;
    };
    OneEList oel[2][HFX_LMAX + 1][HFX_LMAX + 1];
    int *d_sh_atom = nullptr, *d_sh_f0 = nullptr, *d_sh_nprim = nullptr, *d_sh_prim0 = nullptr;
    double *d_alpha = nullptr, *d_coef = nullptr, *d_Z = nullptr;
    double* d_xyz = nullptr;   // 3 nat, geometry
    double* d_grad = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // extra streams (default priority) over which the class launches of a Fock
    // build are dealt, so that the tail of one launch overlaps the next (HFX_STREAMS)
    int nstreams = 3;
    hipStream_t aux[HFX_AUX_STREAMS] = {};
    hipEvent_t evaux[HFX_AUX_STREAMS + 1] = {};
    std::vector<void*> geom_allocs;   // freed on geometry change
};

namespace {

int fail(hfx_t* h, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return 1;
}

#define HIPCHK(h, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(h, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

double seconds_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

int dmalloc(hfx_t* h, void** p, size_t bytes, bool geom) {
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(h, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    if (geom) h->geom_allocs.push_back(*p);
    return 0;
}

// Device buffer of one call (not owned by the handle), freed on every return path.
struct DeviceTemp {
    void* p = nullptr;
    DeviceTemp() = default;
    DeviceTemp(const DeviceTemp&) = delete;
    DeviceTemp& operator=(const DeviceTemp&) = delete;
    ~DeviceTemp() {
        if (p) (void)hipFree(p);
    }
};

template <typename T>
int upload(hfx_t* h, T** dptr, const std::vector<T>& v, bool geom) {
    if (dmalloc(h, (void**)dptr, v.size() * sizeof(T), geom)) return 1;
    if (!v.empty()) HIPCHK(h, hipMemcpy(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
int upload_temp(hfx_t* h, DeviceTemp& b, const std::vector<T>& v) {
    if (dmalloc(h, &b.p, v.size() * sizeof(T), false)) return 1;
    if (!v.empty()) HIPCHK(h, hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

// A non-finite element of an n x n input matrix (column-major) fails the call,
// naming its position: the engine would otherwise spread it over G or g.
int check_finite(hfx_t* h, const double* a, int n, const char* what) {
    const size_t nn = (size_t)n * n;
    for (size_t k = 0; k < nn; k++)
        if (!std::isfinite(a[k]))
            return fail(h, "%s: non-finite element %g at (%zu, %zu)", what, a[k], k % n, k / n);
    return 0;
}

void free_geometry(hfx_t* h) {
    for (void* p : h->geom_allocs) (void)hipFree(p);
    h->geom_allocs.clear();
    h->combos.clear();
    h->rcombos.clear();
    for (int c = 0; c < HFX_NPAIRCLASS; c++)
        for (int w = 0; w < HFX_WIN_MAX; w++) h->ktab[c][w] = hfx_engine::KetTable{};
    h->d_pairs = nullptr;
    h->d_prims = nullptr;
    h->d_primsf = nullptr;
    h->d_pexp = nullptr;
    for (int c = 0; c < HFX_NPAIRCLASS; c++) {
        h->d_cls_list[c] = nullptr;
        h->cls_list[c].clear();
    }
}

// ---------------------------------------------------------------- basis ---
int build_shells(hfx_t* h, int nbf, const int* nprim, const int* istart, const double* x, const double* y,
                 const double* z, const int* l, const int* m, const int* n, int n2, const double* coef,
                 const double* alpha) {
    h->shells.clear();
    h->alpha.clear();
    h->coef.clear();
    h->scale.assign(nbf, 1.0);
    h->bf_atom.assign(nbf, -1);
    int i = 0;
    while (i < nbf) {
        const int L = l[i] + m[i] + n[i];
        if (L < 0 || L > HFX_LMAX) return fail(h, "basis function %d: angular momentum %d not supported (max %d)", i, L, HFX_LMAX);
        const int nc = ncart(L);
        if (i + nc > nbf) return fail(h, "basis function %d: incomplete Cartesian shell", i);
        const int k = nprim[i];
        if (k <= 0) return fail(h, "basis function %d: nprim = %d", i, k);
        for (int c = 0; c < nc; c++) {
            const int f = i + c;
            if (l[f] != cart_exp(L, c, 0) || m[f] != cart_exp(L, c, 1) || n[f] != cart_exp(L, c, 2))
                return fail(h, "basis function %d: powers (%d,%d,%d) out of Cartesian shell order", f, l[f], m[f], n[f]);
            if (nprim[f] != k || x[f] != x[i] || y[f] != y[i] || z[f] != z[i])
                return fail(h, "basis function %d: shell members differ in centre or contraction", f);
            if (istart[f] < 0 || istart[f] + k > n2) return fail(h, "basis function %d: istart out of range", f);
            for (int p = 0; p < k; p++)
                if (alpha[istart[f] + p] != alpha[istart[i] + p])
                    return fail(h, "basis function %d: shell members differ in exponents", f);
        }
        // atom of the shell
        int atom = -1;
        for (int a = 0; a < h->nat; a++)
            if (h->xyz[3 * a] == x[i] && h->xyz[3 * a + 1] == y[i] && h->xyz[3 * a + 2] == z[i]) {
                atom = a;
                break;
            }
        if (atom < 0) return fail(h, "basis function %d: centre is not on any atom", i);
        Shell s;
        s.L = L;
        s.atom = atom;
        s.f0 = i;
        s.nprim = k;
        s.prim0 = (int)h->alpha.size();
        for (int p = 0; p < k; p++) {
            h->alpha.push_back(alpha[istart[i] + p]);
            h->coef.push_back(coef[istart[i] + p]);
        }
        // per-function scale: coef_f = scale_f * coef_shell (checked on every primitive)
        for (int c = 0; c < nc; c++) {
            const int f = i + c;
            const double c0 = coef[istart[i]];
            const double sc = coef[istart[f]] / c0;
            for (int p = 1; p < k; p++) {
                const double r = coef[istart[f] + p] - sc * coef[istart[i] + p];
                if (std::fabs(r) > 1e-12 * std::fabs(coef[istart[f] + p]))
                    return fail(h, "basis function %d: contraction not proportional within its shell", f);
            }
            h->scale[f] = sc;
            h->bf_atom[f] = atom;
        }
        h->shells.push_back(s);
        i += nc;
    }
    h->nbf = nbf;
    // column windows per pair class: only for a ket class whose full K rows (no
    // density staging) would not fit in LDS for some bra class of the basis, in
    // ranges of at most about win_cols columns. (Where the rows fit, windows
    // cost more than they save: the D columns a window reaches extend the pair
    // range past its C columns, and the blocks multiply.)
    bool windows = false, haveL[HFX_LMAX + 1] = {};
    for (const Shell& sh : h->shells) haveL[sh.L] = true;
    for (int c = 0; c < HFX_NPAIRCLASS; c++) {
        const int lc = pair_class_la(c), ld = pair_class_lb(c);
        int ncols = 0;
        for (const Shell& sh : h->shells)
            if (sh.L == lc || (sh.L == ld && ld != lc)) ncols += ncart(sh.L);
        bool wide = false;
        for (int cb = c; cb < HFX_NPAIRCLASS && h->win_cols > 0; cb++) {
            if (!haveL[pair_class_la(cb)] || !haveL[pair_class_lb(cb)]) continue;
            for (int prec = 0; prec < 2; prec++) {
                int nrow = 0, nab = 0;
                hfx_row_shape(cb, c, prec, &nrow, &nab);
                wide = wide || hfx_row_lds(nrow, nab, ncols, 1, 0, prec == 1) > h->lds_budget;
            }
        }
        const int W = wide ? std::max(2, (ncols + h->win_cols - 1) / h->win_cols) : 1;
        h->nwin[c] = std::min(W, HFX_WIN_MAX - 1);   // slot nwin: the class's full table
        windows = windows || h->nwin[c] > 1;
    }
    // LDS columns: the functions of the shells of each L, in shell order, or
    // with windows in the Morton order of their atoms (10 bits per axis over
    // the molecule's bounding box), so that a window's pairs stay close in space
    std::vector<int> order(h->shells.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    if (windows) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int a = 0; a < h->nat; a++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], h->xyz[3 * a + k]);
                hi[k] = std::max(hi[k], h->xyz[3 * a + k]);
            }
        std::vector<unsigned long long> key(h->nat);
        for (int a = 0; a < h->nat; a++) {
            unsigned long long m = 0;
            for (int k = 0; k < 3; k++) {
                const double span = hi[k] - lo[k];
                const unsigned q = span > 0 ? (unsigned)std::min(1023.0, (h->xyz[3 * a + k] - lo[k]) / span * 1023.0) : 0u;
                for (int b = 0; b < 10; b++) m |= (unsigned long long)((q >> b) & 1u) << (3 * b + k);
            }
            key[a] = m;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return key[h->shells[x].atom] < key[h->shells[y].atom]; });
    }
    for (int L = 0; L <= HFX_LMAX; L++) h->bfL[L].clear();
    for (int i : order) {
        Shell& sh = h->shells[i];
        sh.cpos = (int)h->bfL[sh.L].size();
        for (int c = 0; c < ncart(sh.L); c++) h->bfL[sh.L].push_back(sh.f0 + c);
    }
    return 0;
}

// --------------------------------------------------------------- pairs ---
// Primitive pairs are screened with the primitive Schwarz estimate of an s-type
// product, q = |K| sqrt(2 pi^{5/2} / (p^2 sqrt(2p))) = sqrt((ab|ab)) for s
// functions: a pair with q * q_max below 1e-8 times the Schwarz threshold (q_max
// over all primitive pairs) is left out. Every shell pair keeps its place in the
// pair list (index i (i + 1) / 2 + j for shells i >= j).
static constexpr double HFX_PRIM_SCREEN = 1e-8;

double prim_bound(double K, double p) {
    return std::fabs(K) * std::sqrt(2.0 * HFX_PI * HFX_PI * std::sqrt(HFX_PI) / (p * p * std::sqrt(2.0 * p)));
}

void build_pairs(hfx_t* h) {
    h->pairs.clear();
    h->prims.clear();
    h->pexp.clear();
    const int ns = (int)h->shells.size();
    // largest primitive bound of the molecule
    double qmax = 0;
    for (int i = 0; i < ns; i++)
        for (int j = 0; j <= i; j++) {
            const Shell& sa = h->shells[i];
            const Shell& sb = h->shells[j];
            const double* ra = &h->xyz[3 * sa.atom];
            const double* rb = &h->xyz[3 * sb.atom];
            const double r2 = (ra[0] - rb[0]) * (ra[0] - rb[0]) + (ra[1] - rb[1]) * (ra[1] - rb[1]) +
                              (ra[2] - rb[2]) * (ra[2] - rb[2]);
            for (int pa = 0; pa < sa.nprim; pa++)
                for (int pb = 0; pb < sb.nprim; pb++) {
                    const double a = h->alpha[sa.prim0 + pa], b = h->alpha[sb.prim0 + pb], p = a + b;
                    const double K = h->coef[sa.prim0 + pa] * h->coef[sb.prim0 + pb] * std::exp(-a * b / p * r2);
                    qmax = std::max(qmax, prim_bound(K, p));
                }
        }
    const double qcut = qmax > 0 ? HFX_PRIM_SCREEN * h->opts.schwarz_threshold / qmax : 0.0;
    for (int i = 0; i < ns; i++)
        for (int j = 0; j <= i; j++) {
            int A = i, B = j;
            if (h->shells[A].L < h->shells[B].L) std::swap(A, B);
            const Shell& sa = h->shells[A];
            const Shell& sb = h->shells[B];
            const double* ra = &h->xyz[3 * sa.atom];
            const double* rb = &h->xyz[3 * sb.atom];
            ShellPair sp;
            sp.A = A;
            sp.B = B;
            sp.fa = sa.f0;
            sp.fb = sb.f0;
            sp.cls = pair_class(sa.L, sb.L);
            sp.cpa = h->shells[A].cpos;
            sp.cpb = h->shells[B].cpos;
            sp.pad[0] = sp.pad[1] = sp.pad[2] = 0;
            sp.ABx = ra[0] - rb[0];
            sp.ABy = ra[1] - rb[1];
            sp.ABz = ra[2] - rb[2];
            sp.Q = 0;
            sp.p0 = (int)h->prims.size();
            const double r2 = sp.ABx * sp.ABx + sp.ABy * sp.ABy + sp.ABz * sp.ABz;
            for (int pa = 0; pa < sa.nprim; pa++)
                for (int pb = 0; pb < sb.nprim; pb++) {
                    const double a = h->alpha[sa.prim0 + pa], b = h->alpha[sb.prim0 + pb];
                    const double p = a + b;
                    PrimPair pp;
                    pp.p = p;
                    pp.ip = 1.0 / p;
                    pp.Px = (a * ra[0] + b * rb[0]) / p;
                    pp.Py = (a * ra[1] + b * rb[1]) / p;
                    pp.Pz = (a * ra[2] + b * rb[2]) / p;
                    pp.PAx = pp.Px - ra[0];
                    pp.PAy = pp.Py - ra[1];
                    pp.PAz = pp.Pz - ra[2];
                    pp.K = h->coef[sa.prim0 + pa] * h->coef[sb.prim0 + pb] * std::exp(-a * b / p * r2);
                    pp.Kp = HFX_TWO_PI_52 * pp.K / p;
                    if (prim_bound(pp.K, p) < qcut) continue;
                    h->prims.push_back(pp);
                    h->pexp.push_back(PrimExp{a, b});
                }
            sp.np = (int)h->prims.size() - sp.p0;   // 0: the pair drops out through Q = 0
            h->pairs.push_back(sp);
        }
}

// FLOPs of one primitive quartet / one contracted integral of a class
// combination: the operation count of the un-sliced Rys algorithm of
// eri_rys.h (frozen model used for the roofline figure).
struct FlopModel {
    double per_prim;
    double per_int;
};

FlopModel flop_model(int la, int lb, int lc, int ld) {
    const int LAB = la + lb, LCD = lc + ld, NR = (LAB + LCD) / 2 + 1;
    const int NA = ncart(la), NB = ncart(lb), NC = ncart(lc), ND = ncart(ld);
    double f = 0;
    f += 30;                                   // P-Q, T, rho, prefactor (2 div, 1 sqrt)
    f += 2.0 * NR * (3.0 * 13 + 3.0) + 8;      // u_i, w_i as a degree-13 series (frozen model; the
                                               // tables' own degree changes nothing here)
    double per_root = 16;                      // B00, B10, B01, C00, C00' (x3)
    double g = 0;
    g += (LAB > 0 ? 1 : 0) + 3.0 * std::max(0, LAB - 1);
    for (int mm = 0; mm < LCD; mm++)
        for (int nn = 0; nn <= LAB; nn++) g += 1 + (mm > 0 ? 3 : 0) + (nn > 0 ? 3 : 0);
    for (int ib = 1; ib <= lb; ib++) g += 2.0 * (LAB - ib + 1) * (LCD + 1);
    for (int id = 1; id <= ld; id++) g += 2.0 * (la + 1) * (lb + 1) * (LCD - id + 1);
    per_root += 3 * g + 1;
    per_root += 3.0 * NA * NB * NC * ND;       // Ix*Iy, fma with Iz
    f += NR * per_root;
    return FlopModel{f, 12.0};
}

// FLOPs of one primitive quartet / one contracted integral of the analytic
// gradient (grad_rys.h, un-sliced; frozen model of the derivative kernel's
// roofline figure): the energy's set-up and Rys rule for NR = (L + 1)/2 + 1
// roots; per root the three directions' 2-D tables of the combined bra pass
// (built one power higher on A, transferred to B one power higher) and of the
// C pass (one power higher on C), and per integral the nine derivative
// contractions (3 centres x 3 directions: a two-term derivative factor, 3,
// and gam * dI_x * I_y * I_z, 4); per integral once the two-particle density
// factor gam = s (4 D_ab D_cd - D_ac D_bd - D_ad D_bc), 7.
FlopModel grad_flop_model(int la, int lb, int lc, int ld) {
    const int NR = (la + lb + lc + ld + 1) / 2 + 1;
    const double N = (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    auto tables = [](int lab, int lcd, int lak, int lbk, int ldk) {
        double g = (lab > 0 ? 1 : 0) + 3.0 * std::max(0, lab - 1);
        for (int mm = 0; mm < lcd; mm++)
            for (int nn = 0; nn <= lab; nn++) g += 1 + (mm > 0 ? 3 : 0) + (nn > 0 ? 3 : 0);
        for (int ib = 1; ib <= lbk; ib++) g += 2.0 * (lab - ib + 1) * (lcd + 1);
        for (int id = 1; id <= ldk; id++) g += 2.0 * (lak + 1) * (lbk + 1) * (lcd - id + 1);
        return g;
    };
    double f = 30 + 2.0 * NR * (3.0 * 13 + 3.0) + 8;
    double per_root = 16;
    per_root += 3 * tables(la + lb + 1, lc + ld, la + 1, lb + 1, ld);   // A and B pass
    per_root += 3 * tables(la + lb, lc + ld + 1, la, lb, ld);           // C pass
    per_root += 9.0 * 7.0 * N;
    f += NR * per_root;
    return FlopModel{f, 7.0};
}

// ------------------------------------------------------------- screening ---
int compute_schwarz(hfx_t* h) {
    const int np = (int)h->pairs.size();
    std::vector<int> all[HFX_NPAIRCLASS];
    for (int i = 0; i < np; i++) all[h->pairs[i].cls].push_back(i);
    unsigned long long* d_q2 = nullptr;
    if (dmalloc(h, (void**)&d_q2, np * sizeof(unsigned long long), false)) return 1;
    HIPCHK(h, hipMemsetAsync(d_q2, 0, np * sizeof(unsigned long long), h->stream));
    std::vector<int*> tmp;
    for (int c = 0; c < HFX_NPAIRCLASS; c++) {
        if (all[c].empty()) continue;
        int* d_list = nullptr;
        if (upload(h, &d_list, all[c], false)) return 1;
        tmp.push_back(d_list);
        HIPCHK(h, hfx_launch_schwarz(c, h->d_pairs, h->d_prims, d_list, (int)all[c].size(), h->d_scale, d_q2, h->stream));
    }
    std::vector<unsigned long long> q2(np);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(q2.data(), d_q2, np * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int* p : tmp) (void)hipFree(p);
    (void)hipFree(d_q2);
    for (int i = 0; i < np; i++) {
        double v;
        std::memcpy(&v, &q2[i], sizeof v);
        h->pairs[i].Q = std::sqrt(v);
    }
    return 0;
}

// Ket positions per block: whole ranges when a launch has at least
// target_blocks row groups, otherwise about target_blocks blocks' worth
// (multiples of 256). Every block stages its rows' density and flushes its K
// rows, so cutting rows into more blocks costs traffic; the concurrent launches
// of the other streams keep the device busy instead.
int block_chunk(double tot, int ngroups, double target_blocks = 128) {
    if (ngroups >= target_blocks) return 1 << 20;
    return std::max(256, (int)std::ceil(tot / target_blocks / 256.0) * 256);
}

// Primitive-pair count of a ket sublist: counts are rounded up to these bucket
// sizes (beyond the last, to a multiple of it) so that a class has few sublists.
int np_bucket(int np) {
    static const int B[] = {1, 3, 6, 9, 14, 18, 27, 36, 54, 72, 100, 144, 196, 256};
    for (int b : B)
        if (np <= b) return b;
    return (np + 255) / 256 * 256;
}

// Ket table of pair class c (see KetTab), filled on the GPU from the pairs in
// (class, bucket) order: subs = its sublists' (bucket, count) in ascending
// bucket order, sub0 = the class's first entry in that order. aux receives the
// per-ket arrays of the range search (bound, class position, primitive counts
// and their prefix sums).
int build_ket_table(hfx_t* h, int c, int w, const std::vector<std::array<int, 2>>& subs, int sub0,
                    const PairOrder& po, KetAux& aux) {
    hfx_engine::KetTable& T = h->ktab[c][w];
    T = hfx_engine::KetTable{};
    KetTab& t = T.t;
    const int S = (int)subs.size();
    t.S = S;
    T.nd = ncart(pair_class_lb(c));
    T.ncomp = ncart(pair_class_la(c)) * T.nd;
    int nk = 0;
    long long ps = 0;
    for (int s = 0; s < S; s++) {
        t.off[s] = nk;
        t.np[s] = subs[s][0];
        t.psoff[s] = ps;
        t.psoff_f[s] = ps;
        nk += subs[s][1];
        ps += (long long)t.np[s] * KETPRIM_FIELDS * (long long)ket_pitch(subs[s][1]);
    }
    t.off[S] = nk;
    t.nk = nk;
    KetOut o{};
    if (dmalloc(h, (void**)&o.info, (size_t)nk * sizeof(KetInfo), true)) return 1;
    if (dmalloc(h, (void**)&o.cd, 3 * (size_t)nk * sizeof(double), true)) return 1;
    if (dmalloc(h, (void**)&o.shcd, (size_t)nk * sizeof(int), true)) return 1;
    if (dmalloc(h, (void**)&o.qk, (size_t)nk * sizeof(float), true)) return 1;
    if (dmalloc(h, (void**)&o.prim, (size_t)ps * sizeof(double), true)) return 1;
    if (dmalloc(h, (void**)&o.primf, (size_t)ps * sizeof(float), true)) return 1;
    if (dmalloc(h, (void**)&aux.q, (size_t)nk * sizeof(double), true)) return 1;
    if (dmalloc(h, (void**)&aux.pos, (size_t)nk * sizeof(int), true)) return 1;
    if (dmalloc(h, (void**)&aux.np, (size_t)(nk + 1) * sizeof(int), true)) return 1;
    if (dmalloc(h, (void**)&aux.cnp, (size_t)(nk + 1) * sizeof(int), true)) return 1;
    if (dmalloc(h, (void**)&t.Dcd, (size_t)T.ncomp * nk * sizeof(double), true)) return 1;
    if (dmalloc(h, (void**)&t.Jcd, (size_t)T.ncomp * nk * sizeof(double), true)) return 1;
    // padding words of the primitive planes (kets beyond a sublist's count) stay zero
    HIPCHK(h, hipMemsetAsync(o.prim, 0, (size_t)ps * sizeof(double), h->stream));
    HIPCHK(h, hipMemsetAsync(o.primf, 0, (size_t)ps * sizeof(float), h->stream));
    HIPCHK(h, hipMemsetAsync(aux.np + nk, 0, sizeof(int), h->stream));
    t.info = o.info;
    t.cd = o.cd;
    t.prim = o.prim;
    t.primf = o.primf;
    t.shcd = o.shcd;
    t.qk = o.qk;
    HIPCHK(h, hfx_launch_ket_fill(t, o, aux, po.sub_order + sub0, h->d_pairs, h->d_prims, h->d_pexp, po.pos, h->stream));
    return 0;
}

// Screened quartet lists of the current geometry, built on the GPU
// (hfx_lists.hip): class-sorted significant pairs, ket tables, per-row ket
// ranges of every class combination with their quartet counts, row-mode block
// lists. The host sizes the allocations from counts it reads back, plans the
// row-mode LDS and, for class combinations outside row mode (kernel_mode = 1,
// LDS rows that do not fit), builds the tile lists from the ranges.
int build_lists(hfx_t* h) {
    const double thr = h->opts.schwarz_threshold;
    const double thr32 = h->opts.fp32_threshold;
    const int np = (int)h->pairs.size();
    for (int c = 0; c < HFX_NPAIRCLASS; c++) {
        h->cls_list[c].clear();
        h->d_cls_list[c] = nullptr;
    }
    h->combos.clear();
    h->rcombos.clear();
    h->npairs_sig = 0;
    if (np == 0) return 0;
    // column window of every pair, and the LDS columns of every window (over all
    // pairs of the window, significant or not)
    int* d_pwin = nullptr;
    {
        std::vector<int> pwin(np, 0);
        bool windows = false;
        int lo[HFX_NPAIRCLASS][HFX_WIN_MAX], hi[HFX_NPAIRCLASS][HFX_WIN_MAX];
        int dlo[HFX_NPAIRCLASS][HFX_WIN_MAX], dhi[HFX_NPAIRCLASS][HFX_WIN_MAX];
        for (int c = 0; c < HFX_NPAIRCLASS; c++)
            for (int w = 0; w < HFX_WIN_MAX; w++) {
                lo[c][w] = dlo[c][w] = 1 << 30;
                hi[c][w] = dhi[c][w] = 0;
            }
        for (int i = 0; i < np; i++) {
            const ShellPair& sp = h->pairs[i];
            const int c = sp.cls, W = h->nwin[c];
            if (W == 1) continue;
            windows = true;
            const int lc = pair_class_la(c), ld = pair_class_lb(c);
            const int w = std::min(W - 1, (int)((long long)sp.cpa * W / (long long)h->bfL[lc].size()));
            pwin[i] = w;
            if (ld != lc) {
                lo[c][w] = std::min(lo[c][w], sp.cpa);
                hi[c][w] = std::max(hi[c][w], sp.cpa + ncart(lc));
                dlo[c][w] = std::min(dlo[c][w], sp.cpb);
                dhi[c][w] = std::max(dhi[c][w], sp.cpb + ncart(ld));
            } else {
                lo[c][w] = std::min(lo[c][w], std::min(sp.cpa, sp.cpb));
                hi[c][w] = std::max(hi[c][w], std::max(sp.cpa, sp.cpb) + ncart(lc));
            }
        }
        for (int c = 0; c < HFX_NPAIRCLASS; c++) {
            const int lc = pair_class_la(c), ld = pair_class_lb(c);
            for (int w = 0; w <= h->nwin[c] && w < HFX_WIN_MAX; w++) {
                if (h->nwin[c] == 1 || w == h->nwin[c]) {   // all columns
                    h->win_c0[c][w] = 0;
                    h->win_cn[c][w] = (int)h->bfL[lc].size();
                    h->win_d0[c][w] = 0;
                    h->win_dn[c][w] = ld != lc ? (int)h->bfL[ld].size() : 0;
                    continue;
                }
                h->win_c0[c][w] = hi[c][w] > lo[c][w] ? lo[c][w] : 0;
                h->win_cn[c][w] = std::max(0, hi[c][w] - lo[c][w]);
                h->win_d0[c][w] = dhi[c][w] > dlo[c][w] ? dlo[c][w] : 0;
                h->win_dn[c][w] = std::max(0, dhi[c][w] - dlo[c][w]);
            }
        }
        if (windows && upload(h, &d_pwin, pwin, true)) return 1;
    }
    PairOrder po;
    if (dmalloc(h, (void**)&po.cls_order, (size_t)np * sizeof(int), true)) return 1;
    if (dmalloc(h, (void**)&po.sub_order, (size_t)np * sizeof(int), true)) return 1;
    if (dmalloc(h, (void**)&po.pos, (size_t)np * sizeof(int), true)) return 1;
    HIPCHK(h, hfx_pair_order(h->d_pairs, d_pwin, np, thr, po, h->stream));
    h->npairs_sig = po.nsig;
    // with windows also the (class, bucket) order of the full tables
    PairOrder pof;
    if (d_pwin) {
        if (dmalloc(h, (void**)&pof.cls_order, (size_t)np * sizeof(int), true)) return 1;
        if (dmalloc(h, (void**)&pof.sub_order, (size_t)np * sizeof(int), true)) return 1;
        if (dmalloc(h, (void**)&pof.pos, (size_t)np * sizeof(int), true)) return 1;
        HIPCHK(h, hfx_pair_order(h->d_pairs, nullptr, np, thr, pof, h->stream));
    }
    // class lists: views of the sorted array on the device, a host copy for the counts and the tile lists
    std::vector<int> order(po.nsig), hsub, hsubF;
    if (po.nsig > 0)
        HIPCHK(h, hipMemcpy(order.data(), po.cls_order, (size_t)po.nsig * sizeof(int), hipMemcpyDeviceToHost));
    int cls_off[HFX_NPAIRCLASS + 1] = {0};
    for (int c = 0; c < HFX_NPAIRCLASS; c++) {
        cls_off[c + 1] = cls_off[c] + po.ncls[c];
        h->cls_list[c].assign(order.begin() + cls_off[c], order.begin() + cls_off[c + 1]);
        if (po.ncls[c] > 0) h->d_cls_list[c] = po.cls_order + cls_off[c];
    }
    // sublists (bucket, count) of every class and window, in the (class, window, bucket) order
    std::vector<std::array<int, 2>> subs[HFX_NPAIRCLASS][HFX_WIN_MAX];
    int sub0[HFX_NPAIRCLASS][HFX_WIN_MAX] = {};
    {
        int off = 0;
        for (const auto& r : po.runs) {
            if (subs[r[0]][r[1]].empty()) sub0[r[0]][r[1]] = off;
            subs[r[0]][r[1]].push_back({r[2], r[3]});
            off += r[3];
        }
        // full tables of the windowed classes in slot nwin (offsets into pof)
        off = 0;
        for (const auto& r : pof.runs) {
            const int c = r[0], W = h->nwin[c];
            if (W > 1) {
                if (subs[c][W].empty()) sub0[c][W] = off;
                subs[c][W].push_back({r[2], r[3]});
            }
            off += r[3];
        }
    }
    KetAux aux[HFX_NPAIRCLASS][HFX_WIN_MAX] = {};
    const auto tk = std::chrono::steady_clock::now();
    for (int c = 0; c < HFX_NPAIRCLASS; c++)
        for (int w = 0; w <= h->nwin[c] && w < HFX_WIN_MAX; w++) {
            if (subs[c][w].empty()) continue;
            if ((int)subs[c][w].size() > HFX_SUB_MAX)
                return fail(h, "pair class %d has %zu primitive-count sublists (at most %d)", c, subs[c][w].size(),
                            HFX_SUB_MAX);
            const bool full = h->nwin[c] > 1 && w == h->nwin[c];
            if (build_ket_table(h, c, w, subs[c][w], sub0[c][w], full ? pof : po, aux[c][w])) return 1;
        }
    if (h->opts.verbose) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        fprintf(stderr, "hfx: class lists and ket tables %.3f s\n", seconds_since(tk));
    }
    // per class combination: row-mode plan, then the ket ranges of every bra row on the GPU
    struct Pending {
        int cb, ck, w, nb, S, R, dlds;
        bool rowmode;
        bool gonly;           // blocks for the derivative launches only (no J/K, no statistics)
        size_t stat;          // first entry of its statistics
        int* rkb[2];
        int* rke[2];
        hfx_engine::RowCombo rc;
    };
    std::vector<Pending> pend;
    size_t nstat = 0;
    // one pending launch set per ket-table slot: slot w < nwin of a windowed
    // class, slot nwin its full table, slot 0 the table of a class without windows
    auto plan = [&](int cb, int ck, int w, Pending& q) {
        q = Pending{};
        q.cb = cb;
        q.ck = ck;
        q.w = w;
        q.nb = po.ncls[cb];
        q.S = (int)subs[ck][w].size();
        hfx_engine::RowCombo& rc = q.rc;
        rc.cb = cb;
        rc.ck = ck;
        rc.win = w;
        const int lc = pair_class_la(ck), ld = pair_class_lb(ck);
        // LDS columns: the window's (all columns without windows)
        rc.ncols = h->win_cn[ck][w] + (ld != lc ? h->win_dn[ck][w] : 0);
        rc.offd = ld != lc ? h->win_cn[ck][w] : 0;
        rc.col0_c = h->win_c0[ck][w];
        rc.col0_d = ld != lc ? h->win_d0[ck][w] : h->win_c0[ck][w];
        // the FP64 and FP32 launches slice differently and decide the staging of
        // the density rows each within its own budget (dlds bit 0: the FP64
        // launch, bit 1: the FP32 launch); the blocks are planned for the
        // larger LDS of the two
        int nrow64 = 0, nab64 = 0, nrow32 = 0, nab32 = 0;
        hfx_row_shape(cb, ck, 0, &nrow64, &nab64);
        hfx_row_shape(cb, ck, 1, &nrow32, &nab32);
        auto staging = [&]() {
            int m = 0;
            if (h->row_dlds) {
                if (hfx_row_lds(nrow64, nab64, rc.ncols, 1, 1, false) <= std::min(h->lds_budget, h->dlds_budget))
                    m |= 1;
                if (hfx_row_lds(nrow32, nab32, rc.ncols, 1, 1, true) <= std::min(h->lds_budget, h->dlds_budget32))
                    m |= 2;
            }
            return m;
        };
        auto row_lds = [&](int R, int m) {
            return std::max(hfx_row_lds(nrow64, nab64, rc.ncols, R, m & 1, false),
                            hfx_row_lds(nrow32, nab32, rc.ncols, R, (m >> 1) & 1, true));
        };
        int dlds = staging();
        const size_t need = row_lds(1, dlds);
        // rows per block: as many (up to row_rmax) as stay within row_pair_bytes of LDS
        int R = (int)std::min<size_t>(h->row_rmax, h->row_pair_bytes / need);
        if (R == 0 && need <= h->lds_budget) R = 1;
        if (R == 0 && h->row_narrow) {
            // narrow launches: one K row of the a-shell per bra, one row per block
            hfx_row_shape(cb, ck, 4, &nrow64, &nab64);
            hfx_row_shape(cb, ck, 5, &nrow32, &nab32);
            dlds = staging();
            if (row_lds(1, dlds) <= h->lds_budget) {
                R = 1;
                rc.narrow = true;
            }
        }
        q.rowmode = h->opts.kernel_mode == 0 && R >= 1;
        q.R = R;
        q.dlds = dlds;
        rc.R = R;
        rc.dlds = dlds;
    };
    auto push = [&](Pending& q) -> int {
        for (int prec = 0; prec < 2; prec++) {
            if (dmalloc(h, (void**)&q.rkb[prec], (size_t)q.nb * q.S * sizeof(int), true)) return 1;
            if (dmalloc(h, (void**)&q.rke[prec], (size_t)q.nb * q.S * sizeof(int), true)) return 1;
            q.rc.d_kbeg[prec] = q.rkb[prec];
            q.rc.d_kend[prec] = q.rke[prec];
        }
        q.stat = nstat;
        nstat += (size_t)q.S * 4;
        pend.push_back(q);
        return 0;
    };
    for (int c = 0; c < HFX_NPAIRCLASS; c++)
        if ((int)h->bfL[pair_class_la(c)].size() >= 32768 || (int)h->bfL[pair_class_lb(c)].size() >= 32768)
            return fail(h, "ket table: too many functions of one angular momentum");
    for (int cb = 0; cb < HFX_NPAIRCLASS; cb++)
        for (int ck = 0; ck <= cb; ck++) {
            if (po.ncls[cb] == 0 || po.ncls[ck] == 0) continue;
            const int W = h->nwin[ck];
            Pending q;
            if (W == 1) {
                plan(cb, ck, 0, q);
                if (push(q)) return 1;
                continue;
            }
            // windowed class: the windows for J/K (measured faster than the full
            // table for every combination of a class whose rows outgrow the LDS
            // somewhere) and the full table for the derivative launches, which
            // hold no K rows; tiles (kernel mode 1) over the full table
            plan(cb, ck, W, q);
            if (h->opts.kernel_mode == 1) {
                if (push(q)) return 1;
                continue;
            }
            for (int w = 0; w < W; w++) {
                if (subs[ck][w].empty()) continue;
                Pending qw;
                plan(cb, ck, w, qw);
                qw.rc.grad = false;
                if (push(qw)) return 1;
            }
            q.rowmode = true;   // gradient only: row blocks over the full table (no K rows)
            q.gonly = true;
            q.R = q.rc.R = 1;
            q.dlds = q.rc.dlds = 0;
            q.rc.narrow = false;
            q.rc.jk = false;
            if (push(q)) return 1;
        }
    std::vector<double> stats(nstat), sums(pend.size() * 4, 0.0);
    DeviceTemp d_stats, d_sums;
    if (upload_temp(h, d_stats, stats) || upload_temp(h, d_sums, sums)) return 1;
    for (Pending& q : pend) {
        RangeArgs a{};
        a.pairs = h->d_pairs;
        a.bra_list = h->d_cls_list[q.cb];
        a.kt = h->ktab[q.ck][q.w].t;
        a.x = aux[q.ck][q.w];
        a.nb = q.nb;
        a.S = q.S;
        a.same = q.cb == q.ck;
        a.thr = thr;
        a.thr32 = thr32;
        for (int prec = 0; prec < 2; prec++) {
            a.rkb[prec] = q.rkb[prec];
            a.rke[prec] = q.rke[prec];
        }
        a.stats = (double*)d_stats.p + q.stat;
        HIPCHK(h, hfx_launch_row_ranges(a, h->stream));
    }
    // summed spans and non-empty row groups of every row-mode launch (they choose its chunk)
    for (size_t i = 0; i < pend.size(); i++) {
        if (!pend[i].rowmode) continue;
        for (int prec = 0; prec < 2; prec++) {
            const BlockArgsL b{pend[i].rkb[prec], pend[i].rke[prec], pend[i].nb, pend[i].S, pend[i].R, 0};
            HIPCHK(h, hfx_launch_group_spans(b, (double*)d_sums.p + 4 * i + 2 * prec, h->stream));
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(stats.data(), d_stats.p, nstat * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(sums.data(), d_sums.p, sums.size() * sizeof(double), hipMemcpyDeviceToHost));
    bool any_tiles = false;
    for (const Pending& q : pend) any_tiles = any_tiles || !q.rowmode;
    if (any_tiles) {
        hsub.resize(po.nsig);
        if (po.nsig > 0)
            HIPCHK(h, hipMemcpy(hsub.data(), po.sub_order, (size_t)po.nsig * sizeof(int), hipMemcpyDeviceToHost));
        if (d_pwin) {
            hsubF.resize(pof.nsig);
            if (pof.nsig > 0)
                HIPCHK(h, hipMemcpy(hsubF.data(), pof.sub_order, (size_t)pof.nsig * sizeof(int),
                                    hipMemcpyDeviceToHost));
        }
    }
    for (size_t i = 0; i < pend.size(); i++) {
        Pending& q = pend[i];
        const int cb = q.cb, ck = q.ck, w = q.w, nb = q.nb, S = q.S;
        std::vector<int> hkb[2], hke[2];
        if (!q.rowmode)
            for (int prec = 0; prec < 2; prec++) {
                hkb[prec].resize((size_t)nb * S);
                hke[prec].resize((size_t)nb * S);
                HIPCHK(h, hipMemcpy(hkb[prec].data(), q.rkb[prec], hkb[prec].size() * sizeof(int), hipMemcpyDeviceToHost));
                HIPCHK(h, hipMemcpy(hke[prec].data(), q.rke[prec], hke[prec].size() * sizeof(int), hipMemcpyDeviceToHost));
            }
        // the launches' statistics and tile lists (none for the derivative-only blocks)
        for (int s = 0; s < S && !q.gonly; s++) {
            hfx_engine::Combo cmb;
            cmb.cb = cb;
            cmb.ck = ck;
            cmb.win = w;
            cmb.nket = subs[ck][w][s][1];
            cmb.npk = subs[ck][w][s][0];
            cmb.rowmode = q.rowmode;
            const double* st = stats.data() + q.stat + 4 * (size_t)s;
            cmb.nq[0] = st[0];
            cmb.nprimq[0] = st[1];
            cmb.nq[1] = st[2];
            cmb.nprimq[1] = st[3];
            if (q.rowmode) {
                for (int prec = 0; prec < 2; prec++) cmb.ntiles[prec] = cmb.nq[prec] > 0 ? 1 : 0;   // statistics only
            } else {
                // tile mode: the sublist's ket list, per-row ranges relative to it, tiles, gradient blocks
                const int soff = h->ktab[ck][w].t.off[s];
                const std::vector<int>& hs = (h->nwin[ck] > 1 && w == h->nwin[ck]) ? hsubF : hsub;
                std::vector<int> kl(hs.begin() + sub0[ck][w] + soff, hs.begin() + sub0[ck][w] + soff + cmb.nket);
                if (upload(h, &cmb.d_ket, kl, true)) return 1;
                for (int prec = 0; prec < 2; prec++) {
                    std::vector<int> kb(nb), ke(nb), ti0, tj0;
                    for (int r = 0; r < nb; r++) {
                        kb[r] = hkb[prec][(size_t)r * S + s] - soff;
                        ke[r] = hke[prec][(size_t)r * S + s] - soff;
                    }
                    for (int r0 = 0; r0 < nb; r0 += HFX_TILE_ROWS) {
                        int lo = 1 << 30, hi = 0;
                        for (int r = r0; r < std::min(nb, r0 + HFX_TILE_ROWS); r++)
                            if (ke[r] > kb[r]) {
                                lo = std::min(lo, kb[r]);
                                hi = std::max(hi, ke[r]);
                            }
                        if (hi <= lo) continue;
                        for (int j0 = (lo / HFX_TILE_COLS) * HFX_TILE_COLS; j0 < hi; j0 += HFX_TILE_COLS) {
                            ti0.push_back(r0);
                            tj0.push_back(j0);
                        }
                    }
                    cmb.ntiles[prec] = (int)ti0.size();
                    if (cmb.ntiles[prec] > 0) {
                        if (upload(h, &cmb.d_ti0[prec], ti0, true)) return 1;
                        if (upload(h, &cmb.d_tj0[prec], tj0, true)) return 1;
                        if (upload(h, &cmb.d_kbeg[prec], kb, true)) return 1;
                        if (upload(h, &cmb.d_kend[prec], ke, true)) return 1;
                    }
                    // one-row blocks of the fallback gradient kernel
                    std::vector<int> br, bk;
                    double tot = 0;
                    int ngroups = 0;
                    for (int r = 0; r < nb; r++)
                        if (ke[r] > kb[r]) {
                            tot += ke[r] - kb[r];
                            ngroups++;
                        }
                    if (ngroups == 0) continue;
                    const int chunk = block_chunk(tot, ngroups);
                    for (int r = 0; r < nb; r++)
                        for (int k0 = kb[r]; k0 < ke[r]; k0 += chunk) {
                            br.push_back(r);
                            bk.push_back(k0);
                        }
                    cmb.gchunk[prec] = chunk;
                    cmb.gnblk[prec] = (int)br.size();
                    if (upload(h, &cmb.d_gblk_row[prec], br, true)) return 1;
                    if (upload(h, &cmb.d_gblk_k0[prec], bk, true)) return 1;
                }
            }
            if (cmb.ntiles[0] + cmb.ntiles[1] > 0) h->combos.push_back(cmb);
        }
        if (!q.rowmode) continue;
        // row mode: modelled cost (orders the launches), then the blocks of each precision on the GPU
        hfx_engine::RowCombo& rc = q.rc;
        const FlopModel fm = flop_model(pair_class_la(cb), pair_class_lb(cb), pair_class_la(ck), pair_class_lb(ck));
        const double nint = (double)ncart(pair_class_la(cb)) * ncart(pair_class_lb(cb)) * ncart(pair_class_la(ck)) *
                            ncart(pair_class_lb(ck));
        for (int s = 0; s < S; s++)
            for (int prec = 0; prec < 2; prec++) {
                const double* st = stats.data() + q.stat + 4 * (size_t)s + 2 * prec;
                rc.cost[prec] += st[1] * fm.per_prim + st[0] * nint * fm.per_int;
            }
        for (int prec = 0; prec < 2; prec++) {
            const double tot = sums[4 * i + 2 * prec], ngroups = sums[4 * i + 2 * prec + 1];
            if (ngroups == 0) continue;
            const int chunk = block_chunk(tot, (int)ngroups, h->row_blocks);
            const BlockArgsL b{q.rkb[prec], q.rke[prec], nb, S, q.R, chunk};
            HIPCHK(h, hfx_group_blocks(b, &rc.d_blk_row[prec], &rc.d_blk_k0[prec], &rc.nblk[prec], h->geom_allocs,
                                       h->stream));
            rc.chunk[prec] = chunk;
        }
        h->rcombos.push_back(rc);
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->opts.verbose) {
        double q64 = 0, q32 = 0;
        for (auto& c : h->combos) {
            q64 += c.nq[0];
            q32 += c.nq[1];
        }
        fprintf(stderr, "hfx: %d shells, %zu pairs (%.0f significant), quartets fp64 %.4g fp32 %.4g\n",
                (int)h->shells.size(), h->pairs.size(), h->npairs_sig, q64, q32);
        for (auto& c : h->combos)
            fprintf(stderr, "hfx:   class (%d%d|%d%d) ket np %d %s: quartets %.4g / %.4g, prim quartets %.4g / %.4g\n",
                    pair_class_la(c.cb), pair_class_lb(c.cb), pair_class_la(c.ck), pair_class_lb(c.ck),
                    c.npk, c.rowmode ? "rows" : "tiles", c.nq[0], c.nq[1], c.nprimq[0], c.nprimq[1]);
        for (auto& c : h->rcombos)
            fprintf(stderr,
                    "hfx:   rows (%d%d|%d%d) window %d/%d%s: R=%d dlds=%d narrow=%d ncols=%d blocks %d / %d chunk %d / %d\n",
                    pair_class_la(c.cb), pair_class_lb(c.cb), pair_class_la(c.ck), pair_class_lb(c.ck), c.win,
                    h->nwin[c.ck], c.jk ? (c.grad ? "" : " (J/K)") : " (gradient)", c.R, c.dlds, (int)c.narrow,
                    c.ncols, c.nblk[0], c.nblk[1], c.chunk[0], c.chunk[1]);
    }
    return 0;
}

int setup_geometry(hfx_t* h) {
    const auto t0 = std::chrono::steady_clock::now();
    free_geometry(h);
    if (upload(h, &h->d_xyz, h->xyz, true)) return 1;
    build_pairs(h);
    const double t_pairs = seconds_since(t0);
    if (upload(h, &h->d_pairs, h->pairs, true)) return 1;
    if (upload(h, &h->d_prims, h->prims, true)) return 1;
    {
        std::vector<PrimPairF> pf(h->prims.size());
        for (size_t i = 0; i < pf.size(); i++) {
            const PrimPair& q = h->prims[i];
            pf[i] = PrimPairF{(float)q.p, (float)q.ip, (float)q.Px, (float)q.Py, (float)q.Pz,
                              (float)q.PAx, (float)q.PAy, (float)q.PAz, (float)q.K, (float)q.Kp};
        }
        if (upload(h, &h->d_primsf, pf, true)) return 1;
    }
    if (upload(h, &h->d_pexp, h->pexp, true)) return 1;
    const double t_upload = seconds_since(t0);
    if (compute_schwarz(h)) return 1;
    HIPCHK(h, hipMemcpy(h->d_pairs, h->pairs.data(), h->pairs.size() * sizeof(ShellPair), hipMemcpyHostToDevice));
    const double t_schwarz = seconds_since(t0);
    const int rc = build_lists(h);
    if (h->opts.verbose)
        fprintf(stderr, "hfx: geometry set-up %.3f s: pairs %.3f, uploads %.3f, Schwarz bounds %.3f, lists %.3f\n",
                seconds_since(t0), t_pairs, t_upload - t_pairs, t_schwarz - t_upload, seconds_since(t0) - t_schwarz);
    return rc;
}

// ------------------------------------------------------ one-electron ----
// Shell pairs of every (l_A, l_B): unordered (A >= B, the integrals) and
// ordered (the gradient, derivative on the centre of A), geometry-independent.
int build_one_e_lists(hfx_t* h) {
    const int ns = (int)h->shells.size();
    for (int la = 0; la <= HFX_LMAX; la++)
        for (int lb = 0; lb <= HFX_LMAX; lb++)
            for (int ord = 0; ord < 2; ord++) {
                std::vector<int> A, B;
                for (int i = 0; i < ns; i++)
                    for (int j = 0; j < ns; j++) {
                        if (ord == 0 && j > i) continue;
                        if (h->shells[i].L == la && h->shells[j].L == lb) {
                            A.push_back(i);
                            B.push_back(j);
                        }
                    }
                hfx_engine::OneEList& L = h->oel[ord][la][lb];
                L.n = (int)A.size();
                if (L.n == 0) continue;
                if (upload(h, &L.A, A, false) || upload(h, &L.B, B, false)) return 1;
            }
    std::vector<int> sat, sf0, snp, sp0;
    for (const Shell& sh : h->shells) {
        sat.push_back(sh.atom);
        sf0.push_back(sh.f0);
        snp.push_back(sh.nprim);
        sp0.push_back(sh.prim0);
    }
    std::vector<double> Z(h->atno.begin(), h->atno.end());
    if (upload(h, &h->d_sh_atom, sat, false) || upload(h, &h->d_sh_f0, sf0, false) ||
        upload(h, &h->d_sh_nprim, snp, false) || upload(h, &h->d_sh_prim0, sp0, false) ||
        upload(h, &h->d_alpha, h->alpha, false) || upload(h, &h->d_coef, h->coef, false) ||
        upload(h, &h->d_Z, Z, false))
        return 1;
    return 0;
}

OneEArgs one_e_args(hfx_t* h) {
    OneEArgs a{};
    a.sh_atom = h->d_sh_atom;
    a.sh_f0 = h->d_sh_f0;
    a.sh_nprim = h->d_sh_nprim;
    a.sh_prim0 = h->d_sh_prim0;
    a.alpha = h->d_alpha;
    a.coef = h->d_coef;
    a.xyz = h->d_xyz;
    a.Z = h->d_Z;
    a.nat = h->nat;
    a.ztot = h->ztot;
    a.scale = h->d_scale;
    a.n = h->nbf;
    return a;
}

// n x n device work buffers of the host-pointer entry points
int ensure_work(hfx_t* h) {
    const size_t nn = (size_t)h->nbf * h->nbf;
    if (!h->d_P) {
        if (dmalloc(h, (void**)&h->d_P, nn * sizeof(double), false)) return 1;
        if (dmalloc(h, (void**)&h->d_G, nn * sizeof(double), false)) return 1;
        if (dmalloc(h, (void**)&h->d_J, nn * sizeof(double), false)) return 1;
        if (dmalloc(h, (void**)&h->d_K, nn * sizeof(double), false)) return 1;
    }
    return 0;
}

}  // namespace

// =================================================================== ABI ===
extern "C" {

int hfx_default_opts(hfx_opts* o) {
    if (!o) return 1;
    o->schwarz_threshold = 1e-12;
    o->fp32_threshold = 0.0;
    o->device = 0;
    o->verbose = 0;
    o->kernel_mode = 0;
    return 0;
}

int hfx_create(int nbf, const int* nprim, const int* istart, const double* x, const double* y, const double* z,
               const int* l, const int* m, const int* n, int n2, const double* coef, const double* alpha, int nat,
               const int* atno, const double* xyz, const hfx_opts* o, hfx_t** hout) {
    if (!hout) return 1;
    *hout = nullptr;
    hfx_t* h = new hfx_t();
    *hout = h;
    if (o)
        h->opts = *o;
    else
        hfx_default_opts(&h->opts);
    if (nbf <= 0 || !nprim || !istart || !x || !y || !z || !l || !m || !n || !coef || !alpha || nat <= 0 || !atno ||
        !xyz)
        return fail(h, "hfx_create: invalid arguments (nbf=%d nat=%d)", nbf, nat);
    if (!(h->opts.schwarz_threshold >= 0) || !(h->opts.fp32_threshold >= 0))
        return fail(h, "hfx_create: thresholds must be >= 0");
    if (h->opts.kernel_mode < 0 || h->opts.kernel_mode > 1) return fail(h, "hfx_create: kernel_mode must be 0 or 1");
    // row-mode block shape overrides (performance tuning only; results do not depend on them)
    if (const char* e = std::getenv("HFX_ROW_RMAX")) h->row_rmax = std::max(1, std::min(HFX_ROWS_MAX, std::atoi(e)));
    if (const char* e = std::getenv("HFX_ROW_PAIR_KB")) h->row_pair_bytes = (size_t)std::max(1, std::atoi(e)) * 1024;
    if (const char* e = std::getenv("HFX_ROW_DLDS")) h->row_dlds = std::atoi(e) != 0;
    if (const char* e = std::getenv("HFX_ROW_NARROW")) h->row_narrow = std::atoi(e) != 0;
    if (const char* e = std::getenv("HFX_ROW_WINDOW_COLS")) h->win_cols = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("HFX_ROW_DLDS_KB")) h->dlds_budget = (size_t)std::max(1, std::atoi(e)) * 1024;
    if (const char* e = std::getenv("HFX_ROW_DLDS32_KB")) h->dlds_budget32 = (size_t)std::max(1, std::atoi(e)) * 1024;
    if (const char* e = std::getenv("HFX_ROW_BLOCKS")) h->row_blocks = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("HFX_STREAMS")) h->nstreams = std::max(1, std::min(HFX_AUX_STREAMS + 1, std::atoi(e)));
    h->nat = nat;
    h->atno.assign(atno, atno + nat);
    h->ztot = 0;
    for (int a = 0; a < nat; a++) h->ztot += std::fabs((double)atno[a]);
    h->xyz.assign(xyz, xyz + 3 * nat);
    if (build_shells(h, nbf, nprim, istart, x, y, z, l, m, n, n2, coef, alpha)) return 1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(h, "hfx_create: no HIP device available");
    if (h->opts.device < 0 || h->opts.device >= ndev) return fail(h, "hfx_create: device %d out of range", h->opts.device);
    h->device = h->opts.device;
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hfx_scratch_address(h->device, &h->scratch));
    HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    for (int i = 0; i + 1 < h->nstreams; i++) HIPCHK(h, hipStreamCreateWithFlags(&h->aux[i], hipStreamNonBlocking));
    for (int i = 0; i < h->nstreams; i++) HIPCHK(h, hipEventCreateWithFlags(&h->evaux[i], hipEventDisableTiming));
    const size_t nn = (size_t)nbf * nbf;
    if (upload(h, &h->d_scale, h->scale, false)) return 1;
    for (int L = 0; L <= HFX_LMAX; L++)
        if (upload(h, &h->d_bfL[L], h->bfL[L], false)) return 1;
    {
        std::vector<int> sat(h->shells.size());
        for (size_t i = 0; i < h->shells.size(); i++) sat[i] = h->shells[i].atom;
        if (upload(h, &h->d_shell_atom, sat, false)) return 1;
        if (upload(h, &h->d_bf_atom, h->bf_atom, false)) return 1;
        if (dmalloc(h, (void**)&h->d_grad, 3 * (size_t)nat * sizeof(double), false)) return 1;
    }
    if (dmalloc(h, (void**)&h->d_D, nn * sizeof(double), false)) return 1;
    if (dmalloc(h, (void**)&h->d_Jb, nn * sizeof(double), false)) return 1;
    if (dmalloc(h, (void**)&h->d_Kb, nn * sizeof(double), false)) return 1;
    if (build_one_e_lists(h)) return 1;
    return setup_geometry(h);
}

int hfx_set_geometry(hfx_t* h, const double* xyz) {
    if (!h || !xyz) return 1;
    HIPCHK(h, hipSetDevice(h->device));
    h->xyz.assign(xyz, xyz + 3 * h->nat);
    return setup_geometry(h);
}

int hfx_fock_2e_device(hfx_t* h, const double* dP, double* dG, double* dJ, double* dK, hfx_stats* st) {
    if (!h || !dP) return h ? fail(h, "hfx_fock_2e_device: null density") : 1;
    HIPCHK(h, hipSetDevice(h->device));
    auto t0 = std::chrono::steady_clock::now();
    const int n = h->nbf;
    HIPCHK(h, hfx_launch_prep(dP, h->d_scale, h->d_D, h->d_Jb, h->d_Kb, n, h->stream));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    const int nsh = (int)h->shells.size();
    if (h->dthr > 0)
        HIPCHK(h, hfx_launch_shell_dmax(h->d_D, n, h->d_dsh_f0, h->d_dsh_nc, nsh, h->d_dsh, h->stream));
    for (int c = 0; c < HFX_NPAIRCLASS; c++)
        for (int w = 0; w < HFX_WIN_MAX; w++)   // windows and full tables
            if (h->ktab[c][w].t.nk > 0)
                HIPCHK(h, hfx_launch_ket_gather(h->ktab[c][w].t, h->ktab[c][w].ncomp, h->ktab[c][w].nd, h->d_D, n,
                                                h->stream));
    // class launches dealt round-robin over the main and the extra streams,
    // which start after the gathers and are joined before the scatters
    const int ns = h->nstreams;
    if (ns > 1) {
        HIPCHK(h, hipEventRecord(h->evaux[0], h->stream));
        for (int i = 0; i + 1 < ns; i++) HIPCHK(h, hipStreamWaitEvent(h->aux[i], h->evaux[0], 0));
    }
    // largest launches first, each to the stream with the least modelled work so far,
    // so that the small launches at the end fill the tails of the large ones
    std::vector<std::pair<double, int>> order;
    for (size_t i = 0; i < h->rcombos.size(); i++)
        for (int prec = 0; prec < 2; prec++)
            if (h->rcombos[i].jk && h->rcombos[i].nblk[prec] > 0)
                order.push_back({h->rcombos[i].cost[prec], (int)(2 * i + prec)});
    std::stable_sort(order.begin(), order.end(), [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
        return x.first > y.first;
    });
    double load[HFX_AUX_STREAMS + 1] = {};
    for (const auto& o : order) {
        const auto& c = h->rcombos[o.second / 2];
        const int prec = o.second % 2;
        {
            int si = 0;
            for (int i = 1; i < ns; i++)
                if (load[i] < load[si]) si = i;
            load[si] += o.first;
            hipStream_t s = si == 0 ? h->stream : h->aux[si - 1];
            RowArgs r;
            r.pairs = h->d_pairs;
            r.prims = h->d_prims;
            r.primsf = h->d_primsf;
            r.bra_list = h->d_cls_list[c.cb];
            r.kt = h->ktab[c.ck][c.win].t;
            r.kbeg = c.d_kbeg[prec];
            r.kend = c.d_kend[prec];
            r.blk_row = c.d_blk_row[prec];
            r.blk_k0 = c.d_blk_k0[prec];
            r.chunk = c.chunk[prec];
            r.R = c.R;
            r.nbra = (int)h->cls_list[c.cb].size();
            r.ncols = c.ncols;
            r.offd = c.offd;
            r.bf_c = h->d_bfL[pair_class_la(c.ck)];
            r.bf_d = h->d_bfL[pair_class_lb(c.ck)];
            r.col0_c = c.col0_c;
            r.col0_d = c.col0_d;
            r.dlds = ((c.dlds >> prec) & 1) && hfx_row_stage_density(c.cb, c.ck, prec == 1);
            r.D = h->d_D;
            r.J = h->d_Jb;
            r.K = h->d_Kb;
            r.n = n;
            r.dsh = h->dthr > 0 ? h->d_dsh : nullptr;
            r.dthr = (float)h->dthr;
            r.nsh = nsh;
            HIPCHK(h, hfx_launch_jk_rows(c.cb, c.ck, prec == 1, c.narrow, r, c.nblk[prec], s));
        }
    }
    for (int i = 0; i + 1 < ns; i++) {
        HIPCHK(h, hipEventRecord(h->evaux[i + 1], h->aux[i]));
        HIPCHK(h, hipStreamWaitEvent(h->stream, h->evaux[i + 1], 0));
    }
    JKArgs a;
    a.pairs = h->d_pairs;
    a.prims = h->d_prims;
    a.D = h->d_D;
    a.J = h->d_Jb;
    a.K = h->d_Kb;
    a.n = n;
    double q64 = 0, q32 = 0, f64 = 0, f32 = 0, i64 = 0, i32 = 0;
    double cls_fl[HFX_NPAIRCLASS][HFX_NPAIRCLASS][2] = {}, cls_nq[HFX_NPAIRCLASS][HFX_NPAIRCLASS][2] = {},
           cls_np[HFX_NPAIRCLASS][HFX_NPAIRCLASS][2] = {};
    for (const auto& c : h->combos) {
        a.bra_list = h->d_cls_list[c.cb];
        a.ket_list = c.d_ket;
        a.nbra = (int)h->cls_list[c.cb].size();
        const FlopModel fm =
            flop_model(pair_class_la(c.cb), pair_class_lb(c.cb), pair_class_la(c.ck), pair_class_lb(c.ck));
        const double nint = (double)ncart(pair_class_la(c.cb)) * ncart(pair_class_lb(c.cb)) *
                            ncart(pair_class_la(c.ck)) * ncart(pair_class_lb(c.ck));
        for (int prec = 0; prec < 2; prec++) {
            if (c.ntiles[prec] == 0) continue;
            if (!c.rowmode) {
                a.tile_i0 = c.d_ti0[prec];
                a.tile_j0 = c.d_tj0[prec];
                a.kbeg = c.d_kbeg[prec];
                a.kend = c.d_kend[prec];
                HIPCHK(h, hfx_launch_jk(c.cb, c.ck, prec == 1, a, c.ntiles[prec], h->stream));
            }
            const double fl = c.nprimq[prec] * fm.per_prim + c.nq[prec] * nint * fm.per_int;
            cls_fl[c.cb][c.ck][prec] += fl;
            cls_nq[c.cb][c.ck][prec] += c.nq[prec];
            cls_np[c.cb][c.ck][prec] += c.nprimq[prec];
            if (prec == 0) {
                q64 += c.nq[0];
                f64 += fl;
                i64 += c.nq[0] * nint;
            } else {
                q32 += c.nq[1];
                f32 += fl;
                i32 += c.nq[1] * nint;
            }
        }
    }
    for (int c = 0; c < HFX_NPAIRCLASS; c++)
        for (int w = 0; w < HFX_WIN_MAX; w++)   // windows and full tables
            if (h->ktab[c][w].t.nk > 0)
                HIPCHK(h, hfx_launch_ket_scatter(h->ktab[c][w].t, h->ktab[c][w].ncomp, h->ktab[c][w].nd, h->d_Jb, n,
                                                 h->stream));
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hfx_launch_finish(h->d_Jb, h->d_Kb, h->d_scale, dG, dJ, dK, n, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (h->opts.verbose >= 2)   // per class combination and precision: the roofline's FLOP model
        for (int cb = 0; cb < HFX_NPAIRCLASS; cb++)
            for (int ck = 0; ck <= cb; ck++)
                for (int prec = 0; prec < 2; prec++)
                    if (cls_nq[cb][ck][prec] > 0)
                        fprintf(stderr, "hfx: build (%d%d|%d%d) %s quartets %.6g prim_quartets %.6g flops %.6g\n",
                                pair_class_la(cb), pair_class_lb(cb), pair_class_la(ck), pair_class_lb(ck),
                                prec ? "fp32" : "fp64", cls_nq[cb][ck][prec], cls_np[cb][ck][prec],
                                cls_fl[cb][ck][prec]);
    if (st) {
        st->jk_seconds = 1e-3 * ms;
        st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st->quartets_fp64 = q64;
        st->quartets_fp32 = q32;
        st->flops_fp64 = f64;
        st->flops_fp32 = f32;
        st->integrals_fp64 = i64;
        st->integrals_fp32 = i32;
    }
    return 0;
}

int hfx_fock_2e(hfx_t* h, const double* P, double* G, double* J, double* K, hfx_stats* st) {
    if (!h || !P) return h ? fail(h, "hfx_fock_2e: null density") : 1;
    if (check_finite(h, P, h->nbf, "hfx_fock_2e: density")) return 1;
    HIPCHK(h, hipSetDevice(h->device));
    const size_t nn = (size_t)h->nbf * h->nbf;
    if (ensure_work(h)) return 1;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(h, hipMemcpy(h->d_P, P, nn * sizeof(double), hipMemcpyHostToDevice));
    if (hfx_fock_2e_device(h, h->d_P, G ? h->d_G : nullptr, J ? h->d_J : nullptr, K ? h->d_K : nullptr, st)) return 1;
    if (G) HIPCHK(h, hipMemcpy(G, h->d_G, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (J) HIPCHK(h, hipMemcpy(J, h->d_J, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (K) HIPCHK(h, hipMemcpy(K, h->d_K, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (st) st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int hfx_one_electron(hfx_t* h, double* S, double* T, double* V) {
    if (!h) return 1;
    HIPCHK(h, hipSetDevice(h->device));
    if (ensure_work(h)) return 1;
    const size_t nn = (size_t)h->nbf * h->nbf;
    // cleared first; the one-electron lists hold every shell pair (no screening),
    // whose wavefront then stores all elements of its blocks
    HIPCHK(h, hipMemsetAsync(h->d_G, 0, nn * sizeof(double), h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_J, 0, nn * sizeof(double), h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_K, 0, nn * sizeof(double), h->stream));
    OneEArgs a = one_e_args(h);
    a.S = h->d_G;
    a.T = h->d_J;
    a.V = h->d_K;
    for (int la = 0; la <= HFX_LMAX; la++)
        for (int lb = 0; lb <= HFX_LMAX; lb++) {
            const hfx_engine::OneEList& L = h->oel[0][la][lb];
            if (L.n == 0) continue;
            a.pairA = L.A;
            a.pairB = L.B;
            a.npairs = L.n;
            HIPCHK(h, hfx_launch_one_e(la, lb, false, a, h->stream));
        }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (S) HIPCHK(h, hipMemcpy(S, h->d_G, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (T) HIPCHK(h, hipMemcpy(T, h->d_J, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (V) HIPCHK(h, hipMemcpy(V, h->d_K, nn * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int hfx_grad_1e(hfx_t* h, const double* P, const double* W, double* g) {
    if (!h || !P || !W || !g) return h ? fail(h, "hfx_grad_1e: null argument") : 1;
    if (check_finite(h, P, h->nbf, "hfx_grad_1e: density") ||
        check_finite(h, W, h->nbf, "hfx_grad_1e: energy-weighted density"))
        return 1;
    HIPCHK(h, hipSetDevice(h->device));
    if (ensure_work(h)) return 1;
    if (3 * (size_t)h->nat * sizeof(double) > HFX_ROW_LDS_MAX)
        return fail(h, "hfx_grad_1e: %d atoms exceed the per-block accumulator", h->nat);
    const size_t nn = (size_t)h->nbf * h->nbf;
    HIPCHK(h, hipMemcpy(h->d_P, P, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d_G, W, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemsetAsync(h->d_grad, 0, 3 * (size_t)h->nat * sizeof(double), h->stream));
    OneEArgs a = one_e_args(h);
    a.P = h->d_P;
    a.W = h->d_G;
    a.grad = h->d_grad;
    for (int la = 0; la <= HFX_LMAX; la++)
        for (int lb = 0; lb <= HFX_LMAX; lb++) {
            const hfx_engine::OneEList& L = h->oel[1][la][lb];
            if (L.n == 0) continue;
            a.pairA = L.A;
            a.pairB = L.B;
            a.npairs = L.n;
            HIPCHK(h, hfx_launch_one_e(la, lb, true, a, h->stream));
        }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(g, h->d_grad, 3 * (size_t)h->nat * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int hfx_grad_2e(hfx_t* h, const double* P, double* g) {
    if (!h || !P || !g) return h ? fail(h, "hfx_grad_2e: null argument") : 1;
    if (check_finite(h, P, h->nbf, "hfx_grad_2e: density")) return 1;
    HIPCHK(h, hipSetDevice(h->device));
    const int n = h->nbf;
    const size_t nn = (size_t)n * n;
    if (ensure_work(h)) return 1;
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(h, hipMemcpy(h->d_P, P, nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(h, hfx_launch_prep(h->d_P, h->d_scale, h->d_D, h->d_Jb, h->d_Kb, n, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d_grad, 0, 3 * (size_t)h->nat * sizeof(double), h->stream));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    // row mode over the ket tables where the J/K launch is in row mode
    for (int c = 0; c < HFX_NPAIRCLASS; c++)
        for (int w = 0; w < HFX_WIN_MAX; w++)   // windows and full tables
            if (h->ktab[c][w].t.nk > 0)
                HIPCHK(h, hfx_launch_ket_gather(h->ktab[c][w].t, h->ktab[c][w].ncomp, h->ktab[c][w].nd, h->d_D, n,
                                                h->stream));
    bool done[HFX_NPAIRCLASS][HFX_NPAIRCLASS][HFX_WIN_MAX] = {};
    // the row launches dealt largest-first over the streams, as in the Fock build
    const int ns = h->nstreams;
    if (ns > 1) {
        HIPCHK(h, hipEventRecord(h->evaux[0], h->stream));
        for (int i = 0; i + 1 < ns; i++) HIPCHK(h, hipStreamWaitEvent(h->aux[i], h->evaux[0], 0));
    }
    std::vector<std::pair<double, int>> order;
    for (size_t i = 0; i < h->rcombos.size(); i++)
        for (int prec = 0; prec < 2; prec++)
            if (h->rcombos[i].grad && h->rcombos[i].nblk[prec] > 0)
                order.push_back({h->rcombos[i].cost[prec], (int)(2 * i + prec)});
    std::stable_sort(order.begin(), order.end(), [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
        return x.first > y.first;
    });
    double load[HFX_AUX_STREAMS + 1] = {};
    for (const auto& o : order) {
        const auto& c = h->rcombos[o.second / 2];
        const int prec = o.second % 2;
        int nrow = 0, nab = 0;
        hfx_row_shape(c.cb, c.ck, prec ? 3 : 2, &nrow, &nab);   // the launch's slicing (FP64 or FP32)
        // the derivative launch holds no K rows: it stages the density rows
        // whenever they fit, whatever the J/K launch of the combination does
        int dlds = h->row_dlds;
        if (hfx_grad_row_lds(nrow, c.ncols, c.R, dlds, false, h->nat) > HFX_ROW_LDS_MAX) dlds = 0;
        if (hfx_grad_row_lds(nrow, c.ncols, c.R, dlds, false, h->nat) > HFX_ROW_LDS_MAX) continue;
        int si = 0;
        for (int i = 1; i < ns; i++)
            if (load[i] < load[si]) si = i;
        load[si] += o.first;
        hipStream_t s = si == 0 ? h->stream : h->aux[si - 1];
        GradRowArgs r;
        r.pairs = h->d_pairs;
        r.prims = h->d_prims;
        r.pexp = h->d_pexp;
        r.bra_list = h->d_cls_list[c.cb];
        r.kt = h->ktab[c.ck][c.win].t;
        r.R = c.R;
        r.nbra = (int)h->cls_list[c.cb].size();
        r.ncols = c.ncols;
        r.offd = c.offd;
        r.bf_c = h->d_bfL[pair_class_la(c.ck)];
        r.bf_d = h->d_bfL[pair_class_lb(c.ck)];
        r.col0_c = c.col0_c;
        r.col0_d = c.col0_d;
        r.dlds = dlds;
        r.bf_atom = h->d_bf_atom;
        r.nat = h->nat;
        r.D = h->d_D;
        r.grad = h->d_grad;
        r.n = n;
        r.kbeg = c.d_kbeg[prec];
        r.kend = c.d_kend[prec];
        r.blk_row = c.d_blk_row[prec];
        r.blk_k0 = c.d_blk_k0[prec];
        r.chunk = c.chunk[prec];
        HIPCHK(h, hfx_launch_grad_rows(c.cb, c.ck, prec == 1, r, c.nblk[prec], s));
        done[c.cb][c.ck][c.win] = true;
    }
    for (int i = 0; i + 1 < ns; i++) {
        HIPCHK(h, hipEventRecord(h->evaux[i + 1], h->aux[i]));
        HIPCHK(h, hipStreamWaitEvent(h->stream, h->evaux[i + 1], 0));
    }
    GradArgs a;
    a.pairs = h->d_pairs;
    a.prims = h->d_prims;
    a.pexp = h->d_pexp;
    a.shell_atom = h->d_shell_atom;
    a.nat = h->nat;
    a.D = h->d_D;
    a.grad = h->d_grad;
    a.n = n;
    for (const auto& c : h->combos) {
        // the derivative launches of a windowed class ran over its full table
        if (done[c.cb][c.ck][c.win] || (h->nwin[c.ck] > 1 && done[c.cb][c.ck][h->nwin[c.ck]])) continue;
        if (3 * (size_t)h->nat * sizeof(double) > HFX_GRAD_LDS_MAX)
            return fail(h, "hfx_grad_2e: %d atoms need %zu bytes of LDS in the fallback gradient kernel (at most %zu)",
                        h->nat, 3 * (size_t)h->nat * sizeof(double), HFX_GRAD_LDS_MAX);
        a.bra_list = h->d_cls_list[c.cb];
        a.ket_list = c.d_ket;
        for (int prec = 0; prec < 2; prec++) {
            if (c.gnblk[prec] == 0) continue;
            a.kbeg = c.d_kbeg[prec];
            a.kend = c.d_kend[prec];
            a.blk_row = c.d_gblk_row[prec];
            a.blk_k0 = c.d_gblk_k0[prec];
            a.chunk = c.gchunk[prec];
            HIPCHK(h, hfx_launch_grad(c.cb, c.ck, prec == 1, a, c.gnblk[prec], h->stream));
        }
    }
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(g, h->d_grad, 3 * (size_t)h->nat * sizeof(double), hipMemcpyDeviceToHost));
    // statistics: the same quartets as the Fock build, priced by the derivative model
    hfx_stats st{};
    for (const auto& c : h->combos) {
        const int la = pair_class_la(c.cb), lb = pair_class_lb(c.cb), lc = pair_class_la(c.ck),
                  ld = pair_class_lb(c.ck);
        const FlopModel fm = grad_flop_model(la, lb, lc, ld);
        const double nint = (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
        for (int prec = 0; prec < 2; prec++) {
            const double fl = c.nprimq[prec] * fm.per_prim + c.nq[prec] * nint * fm.per_int;
            (prec == 0 ? st.quartets_fp64 : st.quartets_fp32) += c.nq[prec];
            (prec == 0 ? st.flops_fp64 : st.flops_fp32) += fl;
            (prec == 0 ? st.integrals_fp64 : st.integrals_fp32) += c.nq[prec] * nint;
        }
    }
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    st.jk_seconds = 1e-3 * ms;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    h->grad_stats = st;
    return 0;
}

int hfx_grad_2e_stats(const hfx_t* h, hfx_stats* st) {
    if (!h || !st) return 1;
    *st = h->grad_stats;
    return 0;
}

int hfx_eri_shells(hfx_t* h, int nq, const int* shells, double* out) {
    if (!h || nq < 0 || (nq > 0 && (!shells || !out))) return h ? fail(h, "hfx_eri_shells: invalid arguments") : 1;
    if (nq == 0) return 0;
    HIPCHK(h, hipSetDevice(h->device));
    const int ns = (int)h->shells.size();
    // map each requested quartet onto oriented pairs
    auto pair_index = [&](int i, int j) {  // index of unordered shell pair in h->pairs (built i >= j)
        if (i < j) std::swap(i, j);
        return (long)i * (i + 1) / 2 + j;
    };
    std::vector<int> bra(nq), ket(nq);
    std::vector<long long> off(nq), hoff(nq);
    std::vector<int> perm(nq);  // bit0: bra swapped within pair, bit1: ket swapped, bit2: bra<->ket
    long long tot = 0;
    for (int q = 0; q < nq; q++) {
        int A = shells[4 * q], B = shells[4 * q + 1], C = shells[4 * q + 2], D = shells[4 * q + 3];
        if (A < 0 || B < 0 || C < 0 || D < 0 || A >= ns || B >= ns || C >= ns || D >= ns)
            return fail(h, "hfx_eri_shells: shell index out of range in quartet %d", q);
        int p1 = (int)pair_index(A, B), p2 = (int)pair_index(C, D);
        int pm = 0;
        if (h->pairs[p1].A != A) pm |= 1;
        if (h->pairs[p2].A != C) pm |= 2;
        if (h->pairs[p1].cls < h->pairs[p2].cls) {
            std::swap(p1, p2);
            pm |= 4;
        }
        bra[q] = p1;
        ket[q] = p2;
        perm[q] = pm;
        off[q] = tot;
        hoff[q] = tot;
        tot += (long long)ncart(h->shells[A].L) * ncart(h->shells[B].L) * ncart(h->shells[C].L) *
               ncart(h->shells[D].L);
    }
    DeviceTemp b_bra, b_ket, b_off, b_out, b_q;
    if (upload_temp(h, b_bra, bra) || upload_temp(h, b_ket, ket) || upload_temp(h, b_off, off)) return 1;
    if (dmalloc(h, &b_out.p, tot * sizeof(double), false)) return 1;
    int *d_bra = (int*)b_bra.p, *d_ket = (int*)b_ket.p;
    long long* d_off = (long long*)b_off.p;
    double* d_out = (double*)b_out.p;
    // group by class combination
    std::vector<int> order(nq);
    for (int q = 0; q < nq; q++) order[q] = q;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        const int kx = h->pairs[bra[x]].cls * 8 + h->pairs[ket[x]].cls, ky = h->pairs[bra[y]].cls * 8 + h->pairs[ket[y]].cls;
        return kx < ky;
    });
    if (upload_temp(h, b_q, order)) return 1;
    const int* d_q = (const int*)b_q.p;
    int s = 0;
    while (s < nq) {
        const int cb = h->pairs[bra[order[s]]].cls, ck = h->pairs[ket[order[s]]].cls;
        int e = s;
        while (e < nq && h->pairs[bra[order[e]]].cls == cb && h->pairs[ket[order[e]]].cls == ck) e++;
        BlockArgs ba;
        ba.pairs = h->d_pairs;
        ba.prims = h->d_prims;
        ba.bra = d_bra;
        ba.ket = d_ket;
        ba.off = d_off;
        ba.qidx = d_q + s;
        ba.nq = e - s;
        ba.out = d_out;
        ba.packed = 0;
        ba.scale = nullptr;
        HIPCHK(h, hfx_launch_block(cb, ck, ba, h->stream));
        s = e;
    }
    std::vector<double> tmp(tot);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(tmp.data(), d_out, tot * sizeof(double), hipMemcpyDeviceToHost));
    // undo the orientation: kernel layout is (a' b' | c' d') of the oriented pairs
    for (int q = 0; q < nq; q++) {
        const int A = shells[4 * q], B = shells[4 * q + 1], C = shells[4 * q + 2], D = shells[4 * q + 3];
        const int nA = ncart(h->shells[A].L), nB = ncart(h->shells[B].L), nC = ncart(h->shells[C].L),
                  nD = ncart(h->shells[D].L);
        const ShellPair& pb = h->pairs[bra[q]];
        const ShellPair& pk = h->pairs[ket[q]];
        const int n1 = ncart(h->shells[pb.A].L), n2 = ncart(h->shells[pb.B].L), n3 = ncart(h->shells[pk.A].L),
                  n4 = ncart(h->shells[pk.B].L);
        const double* src = &tmp[off[q]];
        double* dst = out + hoff[q];
        for (int a = 0; a < nA; a++)
            for (int b = 0; b < nB; b++)
                for (int c = 0; c < nC; c++)
                    for (int d = 0; d < nD; d++) {
                        // (ab|cd) in request order -> oriented indices
                        int x1 = a, x2 = b, x3 = c, x4 = d;
                        if (perm[q] & 1) std::swap(x1, x2);
                        if (perm[q] & 2) std::swap(x3, x4);
                        if (perm[q] & 4) {
                            std::swap(x1, x3);
                            std::swap(x2, x4);
                        }
                        const double sc = h->scale[h->shells[A].f0 + a] * h->scale[h->shells[B].f0 + b] *
                                          h->scale[h->shells[C].f0 + c] * h->scale[h->shells[D].f0 + d];
                        dst[((a * nB + b) * nC + c) * nD + d] = sc * src[((x1 * n2 + x2) * n3 + x3) * n4 + x4];
                    }
    }
    return 0;
}

int hfx_eri_shells_noalloc(hfx_t* h, int nq, const int* shells, double* out) {
    if (!h || nq < 0 || (nq > 0 && (!shells || !out)))
        return h ? fail(h, "hfx_eri_shells_noalloc: invalid arguments") : 1;
    if (nq == 0) return 0;
    HIPCHK(h, hipSetDevice(h->device));
    if (!h->scratch) return fail(h, "hfx_eri_shells_noalloc: no scratch address for device %d", h->device);
    const int ns = (int)h->shells.size();
    long long tot = 0;
    for (int q = 0; q < nq; q++) {
        const int* r = shells + 4 * q;
        for (int k = 0; k < 4; k++)
            if (r[k] < 0 || r[k] >= ns) return fail(h, "hfx_eri_shells_noalloc: shell index out of range in quartet %d", q);
        tot += (long long)ncart(h->shells[r[0]].L) * ncart(h->shells[r[1]].L) * ncart(h->shells[r[2]].L) *
               ncart(h->shells[r[3]].L);
    }
    // scratch layout: bra, ket, grouped quartet ids (int), block offsets (int64), integrals
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t o_ket = up((size_t)nq * sizeof(int)), o_q = o_ket + up((size_t)nq * sizeof(int));
    const size_t o_off = o_q + up((size_t)nq * sizeof(int)), o_out = o_off + up((size_t)nq * sizeof(long long));
    const size_t need = o_out + (size_t)tot * sizeof(double);
    if (need > HFX_SCRATCH_BYTES)
        return fail(h, "hfx_eri_shells_noalloc: %d quartets with %lld integrals need %zu bytes of scratch, more than "
                       "its %zu bytes (use hfx_eri_shells)", nq, tot, need, HFX_SCRATCH_BYTES);
    unsigned char* base = (unsigned char*)h->scratch;
    int* d_bra = (int*)base;
    int* d_ket = (int*)(base + o_ket);
    int* d_q = (int*)(base + o_q);
    long long* d_off = (long long*)(base + o_off);
    double* d_out = (double*)(base + o_out);
    // oriented shell pairs of quartet q (as hfx_eri_shells): bit 0 of perm = bra
    // swapped within its pair, bit 1 = ket swapped, bit 2 = bra <-> ket
    auto plan = [&](int q, int& p1, int& p2, int& pm) {
        const int A = shells[4 * q], B = shells[4 * q + 1], C = shells[4 * q + 2], D = shells[4 * q + 3];
        p1 = A >= B ? A * (A + 1) / 2 + B : B * (B + 1) / 2 + A;
        p2 = C >= D ? C * (C + 1) / 2 + D : D * (D + 1) / 2 + C;
        pm = (h->pairs[p1].A != A ? 1 : 0) | (h->pairs[p2].A != C ? 2 : 0);
        if (h->pairs[p1].cls < h->pairs[p2].cls) {
            std::swap(p1, p2);
            pm |= 4;
        }
    };
    // plans staged through fixed host buffers: the call allocates no memory
    constexpr int CH = 512;
    int hb[CH], hk[CH];
    long long ho[CH];
    long long run = 0;
    for (int q0 = 0; q0 < nq; q0 += CH) {
        const int m = std::min(CH, nq - q0);
        for (int t = 0; t < m; t++) {
            const int q = q0 + t;
            int pm;
            plan(q, hb[t], hk[t], pm);
            ho[t] = run;
            const int* r = shells + 4 * q;
            run += (long long)ncart(h->shells[r[0]].L) * ncart(h->shells[r[1]].L) * ncart(h->shells[r[2]].L) *
                   ncart(h->shells[r[3]].L);
        }
        // copies on the engine's stream (ordered with the launches); the staging
        // buffers are reused only after they have been read
        HIPCHK(h, hipMemcpyAsync(d_bra + q0, hb, m * sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(d_ket + q0, hk, m * sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(d_off + q0, ho, m * sizeof(long long), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    // one launch per class combination over its quartets (ids grouped in d_q)
    int pos = 0;
    for (int cb = 0; cb < HFX_NPAIRCLASS; cb++)
        for (int ck = 0; ck <= cb; ck++) {
            const int start = pos;
            int fill = 0;
            for (int q = 0; q < nq; q++) {
                int p1, p2, pm;
                plan(q, p1, p2, pm);
                if (h->pairs[p1].cls != cb || h->pairs[p2].cls != ck) continue;
                hb[fill++] = q;
                if (fill == CH) {
                    HIPCHK(h, hipMemcpyAsync(d_q + pos, hb, fill * sizeof(int), hipMemcpyHostToDevice, h->stream));
                    HIPCHK(h, hipStreamSynchronize(h->stream));
                    pos += fill;
                    fill = 0;
                }
            }
            if (fill > 0) {
                HIPCHK(h, hipMemcpyAsync(d_q + pos, hb, fill * sizeof(int), hipMemcpyHostToDevice, h->stream));
                HIPCHK(h, hipStreamSynchronize(h->stream));
                pos += fill;
            }
            if (pos == start) continue;
            BlockArgs ba;
            ba.pairs = h->d_pairs;
            ba.prims = h->d_prims;
            ba.bra = d_bra;
            ba.ket = d_ket;
            ba.off = d_off;
            ba.qidx = d_q + start;
            ba.nq = pos - start;
            ba.out = d_out;
            ba.packed = 0;
            ba.scale = nullptr;
            HIPCHK(h, hfx_launch_block(cb, ck, ba, h->stream));
        }
    if (pos != nq) return fail(h, "hfx_eri_shells_noalloc: %d of %d quartets planned", pos, nq);
    HIPCHK(h, hipMemcpyAsync(out, d_out, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // undo the orientation in place (block by block) and apply the function scales
    double blk[HFX_MAX_BLOCK];
    long long o = 0;
    for (int q = 0; q < nq; q++) {
        const int A = shells[4 * q], B = shells[4 * q + 1], C = shells[4 * q + 2], D = shells[4 * q + 3];
        const int nA = ncart(h->shells[A].L), nB = ncart(h->shells[B].L), nC = ncart(h->shells[C].L),
                  nD = ncart(h->shells[D].L);
        const int sz = nA * nB * nC * nD;
        int p1, p2, pm;
        plan(q, p1, p2, pm);
        const ShellPair& pb = h->pairs[p1];
        const ShellPair& pk = h->pairs[p2];
        const int n2 = ncart(h->shells[pb.B].L), n3 = ncart(h->shells[pk.A].L), n4 = ncart(h->shells[pk.B].L);
        double* dst = out + o;
        std::memcpy(blk, dst, sz * sizeof(double));
        for (int a = 0; a < nA; a++)
            for (int b = 0; b < nB; b++)
                for (int c = 0; c < nC; c++)
                    for (int d = 0; d < nD; d++) {
                        int x1 = a, x2 = b, x3 = c, x4 = d;
                        if (pm & 1) std::swap(x1, x2);
                        if (pm & 2) std::swap(x3, x4);
                        if (pm & 4) {
                            std::swap(x1, x3);
                            std::swap(x2, x4);
                        }
                        const double sc = h->scale[h->shells[A].f0 + a] * h->scale[h->shells[B].f0 + b] *
                                          h->scale[h->shells[C].f0 + c] * h->scale[h->shells[D].f0 + d];
                        dst[((a * nB + b) * nC + c) * nD + d] = sc * blk[((x1 * n2 + x2) * n3 + x3) * n4 + x4];
                    }
        o += sz;
    }
    return 0;
}

int hfx_transform_int2(hfx_t* h, const double* int2, const double* C, double* moint2, long long size) {
    if (!h || !int2 || !C || !moint2) return h ? fail(h, "hfx_transform_int2: null argument") : 1;
    const long long n = h->nbf, m = n * (n + 1) / 2;
    if (size != m * (m + 1) / 2)
        return fail(h, "hfx_transform_int2: size %lld, expected %lld for %lld functions", size, m * (m + 1) / 2, n);
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n4 = (size_t)n * n * n * n;
    size_t freeb = 0, totb = 0;
    HIPCHK(h, hipMemGetInfo(&freeb, &totb));
    const size_t need = 2 * n4 * sizeof(double) + (size_t)size * sizeof(double) + (size_t)n * n * sizeof(double);
    if (need + (256u << 20) > freeb)
        return fail(h, "hfx_transform_int2: needs %.1f GB of device memory, %.1f GB free", need * 1e-9, freeb * 1e-9);
    if (!h->blas) {
        if (rocblas_create_handle(&h->blas) != rocblas_status_success) return fail(h, "rocblas_create_handle failed");
    }
    if (rocblas_set_stream(h->blas, h->stream) != rocblas_status_success) return fail(h, "rocblas_set_stream failed");
    DeviceTemp bX, bY, bP, bC;
    if (dmalloc(h, &bX.p, n4 * sizeof(double), false) || dmalloc(h, &bY.p, n4 * sizeof(double), false) ||
        dmalloc(h, &bP.p, (size_t)size * sizeof(double), false) ||
        dmalloc(h, &bC.p, (size_t)n * n * sizeof(double), false))
        return 1;
    double *X = (double*)bX.p, *Y = (double*)bY.p, *P = (double*)bP.p, *Cd = (double*)bC.p;
    HIPCHK(h, hipMemcpyAsync(P, int2, (size_t)size * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(Cd, C, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hfx_launch_unpack_int2(P, X, (int)n, h->stream));
    // four quarter transforms as column-major GEMMs (C(mu, i) column-major, the
    // array [a][b][c][d] seen as a matrix whose rows run over d, then (c, d), ...):
    const double one = 1.0, zero = 0.0;
    const rocblas_int in = (rocblas_int)n, in2 = (rocblas_int)(n * n), in3 = (rocblas_int)(n * n * n);
    rocblas_status st;
    //  d -> l: Y[l, abc] = sum_d C[d, l] X[d, abc]          -> [a][b][c][l]
    st = rocblas_dgemm(h->blas, rocblas_operation_transpose, rocblas_operation_none, in, in3, in, &one, Cd, in, X, in,
                       &zero, Y, in);
    //  c -> k: per (a, b): X_ab[l, k] = sum_c Y_ab[l, c] C[c, k] -> [a][b][k][l]
    if (st == rocblas_status_success)
        st = rocblas_dgemm_strided_batched(h->blas, rocblas_operation_none, rocblas_operation_none, in, in, in, &one, Y,
                                           in, (rocblas_stride)n * n, Cd, in, 0, &zero, X, in, (rocblas_stride)n * n,
                                           in2);
    //  b -> j: per a: Y_a[(k, l), j] = sum_b X_a[(k, l), b] C[b, j] -> [a][j][k][l]
    if (st == rocblas_status_success)
        st = rocblas_dgemm_strided_batched(h->blas, rocblas_operation_none, rocblas_operation_none, in2, in, in, &one, X,
                                           in2, (rocblas_stride)n * n * n, Cd, in, 0, &zero, Y, in2,
                                           (rocblas_stride)n * n * n, in);
    //  a -> i: X[(j, k, l), i] = sum_a Y[(j, k, l), a] C[a, i] -> [i][j][k][l]
    if (st == rocblas_status_success)
        st = rocblas_dgemm(h->blas, rocblas_operation_none, rocblas_operation_none, in3, in, in, &one, Y, in3, Cd, in,
                           &zero, X, in3);
    if (st != rocblas_status_success) return fail(h, "hfx_transform_int2: rocBLAS error %d", (int)st);
    HIPCHK(h, hfx_launch_pack_int2(X, P, (int)n, h->stream));
    HIPCHK(h, hipMemcpyAsync(moint2, P, (size_t)size * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int hfx_int2(hfx_t* h, double* int2, long long size) {
    if (!h || !int2) return h ? fail(h, "hfx_int2: null output") : 1;
    const long long n = h->nbf, m = n * (n + 1) / 2, need = m * (m + 1) / 2;
    if (size != need) return fail(h, "hfx_int2: size %lld, expected m(m+1)/2 = %lld for n = %lld", size, need, n);
    HIPCHK(h, hipSetDevice(h->device));
    // every unordered pair of shell pairs once (bra index >= ket index), oriented
    // class(bra) >= class(ket); pairs without primitive pairs contribute zeros
    const int np = (int)h->pairs.size();
    std::vector<int> bra, ket;
    for (int p1 = 0; p1 < np; p1++) {
        if (h->pairs[p1].np == 0) continue;
        for (int p2 = 0; p2 <= p1; p2++) {
            if (h->pairs[p2].np == 0) continue;
            const bool sw = h->pairs[p1].cls < h->pairs[p2].cls;
            bra.push_back(sw ? p2 : p1);
            ket.push_back(sw ? p1 : p2);
        }
    }
    const long long nq = (long long)bra.size();
    if (nq > INT32_MAX) return fail(h, "hfx_int2: %lld shell quartets exceed the explicit-quartet launcher", nq);
    // quartets grouped by class combination
    std::vector<int> order(nq), start(64 + 1, 0);
    for (long long q = 0; q < nq; q++) start[h->pairs[bra[q]].cls * 8 + h->pairs[ket[q]].cls + 1]++;
    for (int c = 0; c < 64; c++) start[c + 1] += start[c];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (long long q = 0; q < nq; q++) order[fill[h->pairs[bra[q]].cls * 8 + h->pairs[ket[q]].cls]++] = (int)q;
    }
    DeviceTemp b_bra, b_ket, b_q, b_out;
    if (upload_temp(h, b_bra, bra) || upload_temp(h, b_ket, ket) || upload_temp(h, b_q, order)) return 1;
    if (dmalloc(h, &b_out.p, (size_t)size * sizeof(double), false)) return 1;
    const int *d_bra = (const int*)b_bra.p, *d_ket = (const int*)b_ket.p, *d_q = (const int*)b_q.p;
    double* d_out = (double*)b_out.p;
    HIPCHK(h, hipMemsetAsync(d_out, 0, (size_t)size * sizeof(double), h->stream));
    for (int cb = 0; cb < HFX_NPAIRCLASS; cb++)
        for (int ck = 0; ck <= cb; ck++) {
            const int s0 = start[cb * 8 + ck], s1 = start[cb * 8 + ck + 1];
            if (s1 <= s0) continue;
            BlockArgs ba;
            ba.pairs = h->d_pairs;
            ba.prims = h->d_prims;
            ba.bra = d_bra;
            ba.ket = d_ket;
            ba.off = nullptr;
            ba.qidx = d_q + s0;
            ba.nq = s1 - s0;
            ba.out = d_out;
            ba.packed = 1;
            ba.scale = h->d_scale;
            HIPCHK(h, hfx_launch_block(cb, ck, ba, h->stream));
        }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(int2, d_out, (size_t)size * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int hfx_shells(const hfx_t* h, int* nshell, int* first_bf, int* L, int* atom) {
    if (!h || !nshell) return 1;
    *nshell = (int)h->shells.size();
    for (size_t i = 0; i < h->shells.size(); i++) {
        if (first_bf) first_bf[i] = h->shells[i].f0;
        if (L) L[i] = h->shells[i].L;
        if (atom) atom[i] = h->shells[i].atom;
    }
    return 0;
}

int hfx_pair_bounds(const hfx_t* h, int* npairs, int* A, int* B, double* Q) {
    if (!h || !npairs) return 1;
    *npairs = (int)h->pairs.size();
    for (size_t i = 0; i < h->pairs.size(); i++) {
        if (A) A[i] = h->pairs[i].A;
        if (B) B[i] = h->pairs[i].B;
        if (Q) Q[i] = h->pairs[i].Q;
    }
    return 0;
}

int hfx_set_density_screening(hfx_t* h, double threshold) {
    if (!h) return 1;
    if (!(threshold >= 0)) return fail(h, "hfx_set_density_screening: threshold must be >= 0");
    HIPCHK(h, hipSetDevice(h->device));
    const int nsh = (int)h->shells.size();
    if (threshold > 0 && nsh >= 65536) return fail(h, "hfx_set_density_screening: more than 65535 shells");
    if (threshold > 0 && !h->d_dsh) {
        std::vector<int> f0(nsh), nc(nsh);
        for (int i = 0; i < nsh; i++) {
            f0[i] = h->shells[i].f0;
            nc[i] = ncart(h->shells[i].L);
        }
        if (upload(h, &h->d_dsh_f0, f0, false) || upload(h, &h->d_dsh_nc, nc, false)) return 1;
        if (dmalloc(h, (void**)&h->d_dsh, (size_t)nsh * nsh * sizeof(float), false)) return 1;
    }
    h->dthr = threshold;
    return 0;
}

int hfx_screening(const hfx_t* h, double* npairs, double* nq64, double* nq32) {
    if (!h) return 1;
    double a = 0, b = 0;
    for (const auto& c : h->combos) {
        a += c.nq[0];
        b += c.nq[1];
    }
    if (npairs) *npairs = h->npairs_sig;
    if (nq64) *nq64 = a;
    if (nq32) *nq32 = b;
    return 0;
}

int hfx_work_model(const hfx_t* h, double* model, double* lower) {
    if (!h) return 1;
    double m[2] = {0, 0}, lo[2] = {0, 0};
    for (const auto& c : h->combos) {
        const int la = pair_class_la(c.cb), lb = pair_class_lb(c.cb), lc = pair_class_la(c.ck),
                  ld = pair_class_lb(c.ck);
        const FlopModel fm = flop_model(la, lb, lc, ld);
        const double nint = (double)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
        const int nroots = (la + lb + lc + ld) / 2 + 1;
        for (int prec = 0; prec < 2; prec++) {
            m[prec] += c.nprimq[prec] * fm.per_prim + c.nq[prec] * nint * fm.per_int;
            lo[prec] += c.nprimq[prec] * 3.0 * nroots * nint;
        }
    }
    for (int prec = 0; prec < 2; prec++) {
        if (model) model[prec] = m[prec];
        if (lower) lower[prec] = lo[prec];
    }
    return 0;
}

const char* hfx_last_error(const hfx_t* h) { return h ? h->err.c_str() : "null handle"; }

void hfx_destroy(hfx_t* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    free_geometry(h);
    for (int L = 0; L <= HFX_LMAX; L++)
        if (h->d_bfL[L]) (void)hipFree(h->d_bfL[L]);
    for (int o = 0; o < 2; o++)
        for (int la = 0; la <= HFX_LMAX; la++)
            for (int lb = 0; lb <= HFX_LMAX; lb++) {
                if (h->oel[o][la][lb].A) (void)hipFree(h->oel[o][la][lb].A);
                if (h->oel[o][la][lb].B) (void)hipFree(h->oel[o][la][lb].B);
            }
    void* bufs[] = {h->d_scale, h->d_D, h->d_Jb, h->d_Kb, h->d_P, h->d_G, h->d_J, h->d_K, h->d_shell_atom, h->d_bf_atom,
                    h->d_grad, h->d_sh_atom, h->d_sh_f0, h->d_sh_nprim, h->d_sh_prim0, h->d_alpha, h->d_coef, h->d_Z,
                    h->d_dsh, h->d_dsh_f0, h->d_dsh_nc};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    for (int i = 0; i <= HFX_AUX_STREAMS; i++)
        if (h->evaux[i]) (void)hipEventDestroy(h->evaux[i]);
    for (int i = 0; i < HFX_AUX_STREAMS; i++)
        if (h->aux[i]) (void)hipStreamDestroy(h->aux[i]);
    if (h->blas) (void)rocblas_destroy_handle(h->blas);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
