// Test harness (CPU only): instantiates the engine's Rys quartet templates
// (hfsolver_amd/csrc/eri_rys.h) for the host, so that the kernel arithmetic
// can be checked against the oracle without a GPU. Not part of the product.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../hfsolver_amd/csrc/grad_rys.h"

#define RYS_TABLE_QUALIFIER static const
namespace chk_tables {
#include "../../hfsolver_amd/csrc/rys_tables.inc"
}
#undef RYS_TABLE_QUALIFIER

namespace {

RysTables tabs() {
    return RysTables{chk_tables::RYS_POLY, chk_tables::RYS_HERM_U, chk_tables::RYS_HERM_W, chk_tables::BOYS_TAB,
                     chk_tables::BOYSF_TAB, chk_tables::RYS_CHEBF, chk_tables::BOYS1_TAB};
}

void make_pair(int la, int lb, const double* A, const double* B, int na, const double* aa, const double* ca, int nb,
               const double* ab, const double* cb, ShellPair& sp, std::vector<PrimPair>& prims,
               std::vector<PrimExp>* pexp = nullptr) {
    sp.A = 0;
    sp.B = 1;
    sp.fa = sp.fb = 0;
    sp.cls = pair_class(la, lb);
    sp.cpa = sp.cpb = 0;
    sp.ABx = A[0] - B[0];
    sp.ABy = A[1] - B[1];
    sp.ABz = A[2] - B[2];
    sp.Q = 0;
    sp.p0 = (int)prims.size();
    const double r2 = sp.ABx * sp.ABx + sp.ABy * sp.ABy + sp.ABz * sp.ABz;
    for (int i = 0; i < na; i++)
        for (int j = 0; j < nb; j++) {
            const double a = aa[i], b = ab[j], p = a + b;
            PrimPair pp;
            pp.p = p;
            pp.ip = 1.0 / p;
            pp.Px = (a * A[0] + b * B[0]) / p;
            pp.Py = (a * A[1] + b * B[1]) / p;
            pp.Pz = (a * A[2] + b * B[2]) / p;
            pp.PAx = pp.Px - A[0];
            pp.PAy = pp.Py - A[1];
            pp.PAz = pp.Pz - A[2];
            pp.K = ca[i] * cb[j] * std::exp(-a * b / p * r2);
            pp.Kp = HFX_TWO_PI_52 * pp.K / p;
            prims.push_back(pp);
            if (pexp) pexp->push_back(PrimExp{a, b});
        }
    sp.np = (int)prims.size() - sp.p0;
}

template <int LA, int LB, int LC, int LD, typename R>
void run(const ShellPair& bra, const ShellPair& ket, const PrimPair* prims, double* out) {
    constexpr bool F32 = sizeof(R) == 4;   // the slicing of the FP32 kernels
    constexpr bool SPA = split_a(LA, LB, LC, LD, F32), SPC = split_c(LA, LB, LC, LD, F32);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    for (int split = 0; split < nsplit(LA, LB, LC, LD, F32); split++) {
        const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
        const int sc = SPC ? split % Q::NC : 0;
        R eri[Q::NOUT];
        eri_quartet<LA, LB, LC, LD, SPA, SPC, R>(bra, ket, prims, sa, sc, tabs(), eri);
        for (int a = 0; a < Q::NAt; a++)
            for (int b = 0; b < Q::NB; b++)
                for (int c = 0; c < Q::NCt; c++)
                    for (int d = 0; d < Q::ND; d++) {
                        const int ga = SPA ? sa : a, gc = SPC ? sc : c;
                        out[((ga * Q::NB + b) * Q::NC + gc) * Q::ND + d] =
                            (double)eri[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d];
                    }
    }
}

template <int LA, int LB, int LC, int LD, typename R>
void run_grad(const ShellPair& bra, const ShellPair& ket, const PrimPair* prims, const PrimExp* pexp,
              const double* gam_full, double* g) {
    constexpr bool G32 = sizeof(R) == 4;
    constexpr bool SPA = gsplit_a(LA, LB, LC, LD, G32), SPC = gsplit_c(LA, LB, LC, LD, G32);
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int split = 0; split < gnsplit(LA, LB, LC, LD, G32); split++) {
        const int sa = SPA ? (SPC ? split / Q::NC : split) : 0;
        const int sc = SPC ? split % Q::NC : 0;
        R gam[Q::NOUT];
        for (int a = 0; a < Q::NAt; a++)
            for (int b = 0; b < Q::NB; b++)
                for (int c = 0; c < Q::NCt; c++)
                    for (int d = 0; d < Q::ND; d++) {
                        const int ga = SPA ? sa : a, gc = SPC ? sc : c;
                        gam[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d] =
                            R(gam_full[((ga * Q::NB + b) * Q::NC + gc) * Q::ND + d]);
                    }
        grad_quartet<LA, LB, LC, LD, SPA, SPC, R>(bra, ket, prims, pexp, sa, sc, tabs(), gam, acc);
    }
    for (int k = 0; k < 9; k++) g[k] = acc[k];
}

template <int X, int Y, typename R>
void run_grad_combo(const ShellPair& bra, const ShellPair& ket, const PrimPair* prims, const PrimExp* pexp,
                    const double* gam, double* g) {
    run_grad<pair_class_la(X), pair_class_lb(X), pair_class_la(Y), pair_class_lb(Y), R>(bra, ket, prims, pexp, gam, g);
}

template <int X, int Y, typename R>
void run_combo(const ShellPair& bra, const ShellPair& ket, const PrimPair* prims, double* out) {
    run<pair_class_la(X), pair_class_lb(X), pair_class_la(Y), pair_class_lb(Y), R>(bra, ket, prims, out);
}

typedef void (*fn)(const ShellPair&, const ShellPair&, const PrimPair*, double*);

#define ROW(X, R)                                                                                       \
    {run_combo<X, 0, R>, X >= 1 ? run_combo<X, (X >= 1 ? 1 : 0), R> : nullptr,                          \
     X >= 2 ? run_combo<X, (X >= 2 ? 2 : 0), R> : nullptr, X >= 3 ? run_combo<X, (X >= 3 ? 3 : 0), R> : nullptr, \
     X >= 4 ? run_combo<X, (X >= 4 ? 4 : 0), R> : nullptr, X >= 5 ? run_combo<X, (X >= 5 ? 5 : 0), R> : nullptr}

typedef void (*gfn)(const ShellPair&, const ShellPair&, const PrimPair*, const PrimExp*, const double*, double*);
#define GROW(X, R)                                                                                             \
    {run_grad_combo<X, 0, R>, X >= 1 ? run_grad_combo<X, (X >= 1 ? 1 : 0), R> : nullptr,                       \
     X >= 2 ? run_grad_combo<X, (X >= 2 ? 2 : 0), R> : nullptr, X >= 3 ? run_grad_combo<X, (X >= 3 ? 3 : 0), R> : nullptr, \
     X >= 4 ? run_grad_combo<X, (X >= 4 ? 4 : 0), R> : nullptr, X >= 5 ? run_grad_combo<X, (X >= 5 ? 5 : 0), R> : nullptr}
const gfn G64[6][6] = {GROW(0, double), GROW(1, double), GROW(2, double), GROW(3, double), GROW(4, double), GROW(5, double)};
const gfn G32[6][6] = {GROW(0, float), GROW(1, float), GROW(2, float), GROW(3, float), GROW(4, float), GROW(5, float)};

const fn T64[6][6] = {ROW(0, double), ROW(1, double), ROW(2, double), ROW(3, double), ROW(4, double), ROW(5, double)};
const fn T32[6][6] = {ROW(0, float), ROW(1, float), ROW(2, float), ROW(3, float), ROW(4, float), ROW(5, float)};

}  // namespace

extern "C" {

// Block (ab|cd) of one shell quartet, a fastest-varying last (d): out[((a*nB+b)*nC+c)*nD+d].
// Requires la >= lb, lc >= ld and class(la,lb) >= class(lc,ld).
int chk_eri_quartet(int la, int lb, int lc, int ld, const double* A, const double* B, const double* C,
                    const double* D, int na, const double* aa, const double* ca, int nb, const double* ab,
                    const double* cb, int nc, const double* ac, const double* cc, int nd, const double* ad,
                    const double* cd, int fp32, double* out) {
    if (la < lb || lc < ld || la > 2 || lc > 2) return 1;
    const int X = pair_class(la, lb), Y = pair_class(lc, ld);
    if (X < Y) return 2;
    std::vector<PrimPair> prims;
    ShellPair bra, ket;
    make_pair(la, lb, A, B, na, aa, ca, nb, ab, cb, bra, prims);
    make_pair(lc, ld, C, D, nc, ac, cc, nd, ad, cd, ket, prims);
    (fp32 ? T32 : T64)[X][Y](bra, ket, prims.data(), out);
    return 0;
}

// g[9] = sum_abcd gam_abcd d(ab|cd)/d(A,B,C) (gam full block, d fastest)
int chk_grad_quartet(int la, int lb, int lc, int ld, const double* A, const double* B, const double* C,
                     const double* D, int na, const double* aa, const double* ca, int nb, const double* ab,
                     const double* cb, int nc, const double* ac, const double* cc, int nd, const double* ad,
                     const double* cd, const double* gam, int fp32, double* g) {
    if (la < lb || lc < ld || la > 2 || lc > 2) return 1;
    const int X = pair_class(la, lb), Y = pair_class(lc, ld);
    if (X < Y) return 2;
    std::vector<PrimPair> prims;
    std::vector<PrimExp> pexp;
    ShellPair bra, ket;
    make_pair(la, lb, A, B, na, aa, ca, nb, ab, cb, bra, prims, &pexp);
    make_pair(lc, ld, C, D, nc, ac, cc, nd, ad, cd, ket, prims, &pexp);
    (fp32 ? G32 : G64)[X][Y](bra, ket, prims.data(), pexp.data(), gam, g);
    return 0;
}

double chk_rys(int nr, double T, double* u, double* w) {
    switch (nr) {
        case -1: rys_roots_small<1>(T, u, w, chk_tables::BOYS_TAB); break;
        case -2: rys_roots_small<2>(T, u, w, chk_tables::BOYS_TAB); break;
        case -21:
        case -22:
        case -23:
        case -24:
        case -25: {  // single-precision Chebyshev rules of the FP32 kernels
            float uf[5], wf[5];
            const float Tf = (float)T;
            switch (nr) {
                case -21: rys_roots_f<1>(Tf, uf, wf, tabs()); break;
                case -22: rys_roots_f<2>(Tf, uf, wf, tabs()); break;
                case -23: rys_roots_f<3>(Tf, uf, wf, tabs()); break;
                case -24: rys_roots_f<4>(Tf, uf, wf, tabs()); break;
                default: rys_roots_f<5>(Tf, uf, wf, tabs()); break;
            }
            for (int i = 0; i < -nr - 20; i++) {
                u[i] = uf[i];
                w[i] = wf[i];
            }
            break;
        }
        case -11:
        case -12: {  // single-precision rules of the FP32 kernels
            float uf[2], wf[2];
            if (nr == -11)
                rys_roots_small_f<1>((float)T, uf, wf, chk_tables::BOYSF_TAB);
            else
                rys_roots_small_f<2>((float)T, uf, wf, chk_tables::BOYSF_TAB);
            for (int i = 0; i < -nr - 10; i++) {
                u[i] = uf[i];
                w[i] = wf[i];
            }
            break;
        }
        case -31: rys_roots_one<BOYS_K>(T, u, w, chk_tables::BOYS1_TAB); break;     // compact one-root table
        case -41: rys_roots_one<BOYS1_KF>(T, u, w, chk_tables::BOYS1_TAB); break;   // ... with the FP32 kernels' terms
        case -51: rys_roots_small<1, HFX_BOYS_K32>(T, u, w, chk_tables::BOYS_TAB); break;   // FP32 kernels' terms
        case -52: rys_roots_small<2, HFX_BOYS_K32>(T, u, w, chk_tables::BOYS_TAB); break;
        case 1: rys_roots<1>(T, u, w, tabs()); break;
        case 2: rys_roots<2>(T, u, w, tabs()); break;
        case 3: rys_roots<3>(T, u, w, tabs()); break;
        case 4: rys_roots<4>(T, u, w, tabs()); break;
        case 5: rys_roots<5>(T, u, w, tabs()); break;
        default: return -1;
    }
    return 0;
}
}
