/*
 * ORACLE - test infrastructure only.
 *
 * A plain-C restatement of hfsolver's closed-shell RHF integral path, used as
 * the checker for the HIP engine. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * What it restates (reference file:line, paths relative to hfsolver's src/):
 *   orc_boys            special_functions.f90:189-248  (Shavitt series + recursions)
 *   orc_overlap         cints.c:212-244               (THO eq. 2.12)
 *   orc_kinetic         cints.c:186-210
 *   orc_nuclear         cints.c:246-307               (THO eq. 2.18, 3.1)
 *   orc_eri_prim        cints.c:103-183               (THO eq. 2.22, B arrays)
 *   orc_eri_contr       cints.c:72-101                (contr_coulomb)
 *   orc_getS/T/V        qc.c:6-93                     (bf x bf loops, col-major)
 *   orc_getInts         qc.c:95-132 + cints.c:327-347 (packed 8-fold int2; 64-bit index)
 *   orc_jk              scf.f90:26-37 / :39-58        (formG: J, K of a density)
 *   orc_eri_bf          one contracted (ij|kl) for arbitrary basis functions
 *   orc_schwarz,        Schwarz bounds of shell pairs and J, K over the
 *   orc_jk_screened     Schwarz-screened unique shell quartets (the direct
 *                       engine's screening rule; same THO integrals as orc_jk)
 *   orc_grad_*          derivative integrals by the centre-derivative identity
 *                       d/dA_x x_A^l e^{-a x_A^2} = 2a x_A^{l+1}(...) - l x_A^{l-1}(...)
 *                       applied to the same THO primitives (no reference
 *                       counterpart: hfsolver has no gradient code).
 *
 * Parity pins: orc_boys against the closed forms of test_fm.f90:46-68 (see
 * tests/test_oracle.py); orc_getS/T/V, orc_getInts and the SCF built on them
 * against the compiled reference's own outputs (tests/golden/) and the
 * reference's literature energies (test_gaussians_molecules.f90).
 *
 * Integer helpers (fact_ratio2, binomial, fact2) return exact integers, as the
 * reference's int-returning versions do for the small arguments used here.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXL 16

static const double PI = 3.14159265358979323846;

static double fact_d(int n) {
    double r = 1.0;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

/* a! / (b! (a-2b)!) */
static double fact_ratio2(int a, int b) { return fact_d(a) / fact_d(b) / fact_d(a - 2 * b); }
static double binom(int a, int b) { return fact_d(a) / (fact_d(b) * fact_d(a - b)); }
static double fact2(int n) { return n <= 1 ? 1.0 : n * fact2(n - 2); }

static double ipow(double x, int n) {
    if (n < 0) return 1.0 / ipow(x, -n);
    double r = 1.0;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

/* ---------------------------------------------------------------- Boys --- */
int orc_boys(int maxm, double t, double *F) {
    if (maxm < 0 || maxm > 500 || t < 0) return 1;
    if (t < maxm + 0.5) {
        double term = 1.0 / (2 * maxm + 1), s = term;
        int m = 1;
        while (term / s > DBL_EPSILON) {
            term = term * (2 * t) / (2 * maxm + 2 * m + 1);
            s = s + term;
            m = m + 1;
        }
        double et = exp(-t);
        F[maxm] = s * et;
        for (m = maxm - 1; m >= 0; m--) F[m] = (2 * t * F[m + 1] + et) / (2 * m + 1);
    } else {
        double et = exp(-t);
        F[0] = 0.5 * sqrt(PI / t) * erf(sqrt(t));
        for (int m = 0; m < maxm; m++) F[m + 1] = ((2 * m + 1) * F[m] - et) / (2 * t);
    }
    return 0;
}

/* ------------------------------------------------------- THO helpers --- */
static double binomial_prefactor(int s, int ia, int ib, double xpa, double xpb) {
    double sum = 0.0;
    for (int t = 0; t < s + 1; t++)
        if ((s - ia <= t) && (t <= ib))
            sum += binom(ia, s - t) * binom(ib, t) * ipow(xpa, ia - s + t) * ipow(xpb, ib - t);
    return sum;
}

static double dist2(double x1, double y1, double z1, double x2, double y2, double z2) {
    return (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2);
}

static double pc1(double a, double xa, double b, double xb) { return (a * xa + b * xb) / (a + b); }

/* ---------------------------------------------------------- overlap ----- */
static double overlap_1d(int l1, int l2, double PAx, double PBx, double gamma) {
    double sum = 0.0;
    for (int i = 0; i < 1 + (l1 + l2) / 2; i++)
        sum += binomial_prefactor(2 * i, l1, l2, PAx, PBx) * fact2(2 * i - 1) / ipow(2 * gamma, i);
    return sum;
}

double orc_overlap(double a1, int l1, int m1, int n1, double xa, double ya, double za,
                   double a2, int l2, int m2, int n2, double xb, double yb, double zb) {
    if (l1 < 0 || m1 < 0 || n1 < 0 || l2 < 0 || m2 < 0 || n2 < 0) return 0.0;
    double rab2 = dist2(xa, ya, za, xb, yb, zb), g = a1 + a2;
    double xp = pc1(a1, xa, a2, xb), yp = pc1(a1, ya, a2, yb), zp = pc1(a1, za, a2, zb);
    double pre = pow(PI / g, 1.5) * exp(-a1 * a2 * rab2 / g);
    return pre * overlap_1d(l1, l2, xp - xa, xp - xb, g) * overlap_1d(m1, m2, yp - ya, yp - yb, g) *
           overlap_1d(n1, n2, zp - za, zp - zb, g);
}

double orc_kinetic(double a1, int l1, int m1, int n1, double xa, double ya, double za,
                   double a2, int l2, int m2, int n2, double xb, double yb, double zb) {
    if (l1 < 0 || m1 < 0 || n1 < 0 || l2 < 0 || m2 < 0 || n2 < 0) return 0.0;
    double t0 = a2 * (2 * (l2 + m2 + n2) + 3) * orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2, m2, n2, xb, yb, zb);
    double t1 = -2 * a2 * a2 *
                (orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2 + 2, m2, n2, xb, yb, zb) +
                 orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2, m2 + 2, n2, xb, yb, zb) +
                 orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2, m2, n2 + 2, xb, yb, zb));
    double t2 = -0.5 * (l2 * (l2 - 1) * orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2 - 2, m2, n2, xb, yb, zb) +
                        m2 * (m2 - 1) * orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2, m2 - 2, n2, xb, yb, zb) +
                        n2 * (n2 - 1) * orc_overlap(a1, l1, m1, n1, xa, ya, za, a2, l2, m2, n2 - 2, xb, yb, zb));
    return t0 + t1 + t2;
}

/* -------------------------------------------------- nuclear attraction --- */
static double A_term(int i, int r, int u, int l1, int l2, double PAx, double PBx, double CPx, double g) {
    return ipow(-1, i) * binomial_prefactor(i, l1, l2, PAx, PBx) * ipow(-1, u) * fact_d(i) *
           ipow(CPx, i - 2 * r - 2 * u) * ipow(0.25 / g, r + u) / fact_d(r) / fact_d(u) / fact_d(i - 2 * r - 2 * u);
}

static void A_array(int l1, int l2, double PA, double PB, double CP, double g, double *A) {
    int Imax = l1 + l2 + 1;
    for (int i = 0; i < Imax; i++) A[i] = 0.0;
    for (int i = 0; i < Imax; i++)
        for (int r = 0; r < i / 2 + 1; r++)
            for (int u = 0; u < (i - 2 * r) / 2 + 1; u++) A[i - 2 * r - u] += A_term(i, r, u, l1, l2, PA, PB, CP, g);
}

double orc_nuclear(double x1, double y1, double z1, int l1, int m1, int n1, double a1,
                   double x2, double y2, double z2, int l2, int m2, int n2, double a2,
                   double x3, double y3, double z3) {
    if (l1 < 0 || m1 < 0 || n1 < 0 || l2 < 0 || m2 < 0 || n2 < 0) return 0.0;
    double g = a1 + a2;
    double xp = pc1(a1, x1, a2, x2), yp = pc1(a1, y1, a2, y2), zp = pc1(a1, z1, a2, z2);
    double rab2 = dist2(x1, y1, z1, x2, y2, z2), rcp2 = dist2(x3, y3, z3, xp, yp, zp);
    double Ax[ORC_MAXL], Ay[ORC_MAXL], Az[ORC_MAXL], F[3 * ORC_MAXL];
    A_array(l1, l2, xp - x1, xp - x2, xp - x3, g, Ax);
    A_array(m1, m2, yp - y1, yp - y2, yp - y3, g, Ay);
    A_array(n1, n2, zp - z1, zp - z2, zp - z3, g, Az);
    orc_boys(l1 + l2 + m1 + m2 + n1 + n2, rcp2 * g, F);
    double sum = 0.0;
    for (int I = 0; I < l1 + l2 + 1; I++)
        for (int J = 0; J < m1 + m2 + 1; J++)
            for (int K = 0; K < n1 + n2 + 1; K++) sum += Ax[I] * Ay[J] * Az[K] * F[I + J + K];
    return -2 * PI / g * exp(-a1 * a2 * rab2 / g) * sum;
}

/* -------------------------------------------------------- ERI (THO) ---- */
static double Bfunc(int i, int r, double g) { return fact_ratio2(i, r) * ipow(4 * g, r - i); }

static double fB(int i, int l1, int l2, double px, double ax, double bx, int r, double g) {
    return binomial_prefactor(i, l1, l2, px - ax, px - bx) * Bfunc(i, r, g);
}

static double B_term(int i1, int i2, int r1, int r2, int u, int l1, int l2, int l3, int l4, double Px, double Ax,
                     double Bx, double Qx, double Cx, double Dx, double g1, double g2, double delta) {
    int e = i1 + i2 - 2 * (r1 + r2);
    return fB(i1, l1, l2, Px, Ax, Bx, r1, g1) * ipow(-1, i2) * fB(i2, l3, l4, Qx, Cx, Dx, r2, g2) * ipow(-1, u) *
           fact_ratio2(e, u) * ipow(Qx - Px, e - 2 * u) / ipow(delta, e - u);
}

static void B_array(int l1, int l2, int l3, int l4, double p, double a, double b, double q, double c, double d,
                    double g1, double g2, double delta, double *B) {
    int Imax = l1 + l2 + l3 + l4 + 1;
    for (int i = 0; i < Imax; i++) B[i] = 0.0;
    for (int i1 = 0; i1 < l1 + l2 + 1; i1++)
        for (int i2 = 0; i2 < l3 + l4 + 1; i2++)
            for (int r1 = 0; r1 < i1 / 2 + 1; r1++)
                for (int r2 = 0; r2 < i2 / 2 + 1; r2++)
                    for (int u = 0; u < (i1 + i2) / 2 - r1 - r2 + 1; u++) {
                        int I = i1 + i2 - 2 * (r1 + r2) - u;
                        B[I] += B_term(i1, i2, r1, r2, u, l1, l2, l3, l4, p, a, b, q, c, d, g1, g2, delta);
                    }
}

double orc_eri_prim(double xa, double ya, double za, int la, int ma, int na, double aa,
                    double xb, double yb, double zb, int lb, int mb, int nb, double ab,
                    double xc, double yc, double zc, int lc, int mc, int nc, double ac,
                    double xd, double yd, double zd, int ld, int md, int nd, double ad) {
    if (la < 0 || ma < 0 || na < 0 || lb < 0 || mb < 0 || nb < 0 || lc < 0 || mc < 0 || nc < 0 || ld < 0 ||
        md < 0 || nd < 0)
        return 0.0;
    double rab2 = dist2(xa, ya, za, xb, yb, zb), rcd2 = dist2(xc, yc, zc, xd, yd, zd);
    double xp = pc1(aa, xa, ab, xb), yp = pc1(aa, ya, ab, yb), zp = pc1(aa, za, ab, zb);
    double xq = pc1(ac, xc, ad, xd), yq = pc1(ac, yc, ad, yd), zq = pc1(ac, zc, ad, zd);
    double rpq2 = dist2(xp, yp, zp, xq, yq, zq);
    double g1 = aa + ab, g2 = ac + ad, delta = (1. / g1 + 1. / g2) / 4.;
    double Bx[4 * ORC_MAXL], By[4 * ORC_MAXL], Bz[4 * ORC_MAXL], F[12 * ORC_MAXL];
    B_array(la, lb, lc, ld, xp, xa, xb, xq, xc, xd, g1, g2, delta, Bx);
    B_array(ma, mb, mc, md, yp, ya, yb, yq, yc, yd, g1, g2, delta, By);
    B_array(na, nb, nc, nd, zp, za, zb, zq, zc, zd, g1, g2, delta, Bz);
    orc_boys(la + lb + lc + ld + ma + mb + mc + md + na + nb + nc + nd, 0.25 * rpq2 / delta, F);
    double sum = 0.0;
    for (int I = 0; I < la + lb + lc + ld + 1; I++)
        for (int J = 0; J < ma + mb + mc + md + 1; J++)
            for (int K = 0; K < na + nb + nc + nd + 1; K++) sum += Bx[I] * By[J] * Bz[K] * F[I + J + K];
    return 2. * pow(PI, 2.5) / (g1 * g2 * sqrt(g1 + g2)) * exp(-aa * ab * rab2 / g1) * exp(-ac * ad * rcd2 / g2) *
           sum;
}

/* ------------------------------------------------- basis description ---- */
/* The per-basis-function SoA arrays that gaussints hands to the C kernels
 * (basis.f90:541-567): nprim, istart (0-based), centres, powers, coef, alpha. */
typedef struct {
    int nbf;
    const int *nprim, *istart;
    const double *x, *y, *z;
    const int *l, *m, *n;
    const double *coef, *alpha;
} orc_basis;

static orc_basis mkbasis(int nbf, const int *nprim, const int *istart, const double *x, const double *y,
                         const double *z, const int *l, const int *m, const int *n, const double *coef,
                         const double *alpha) {
    orc_basis b = {nbf, nprim, istart, x, y, z, l, m, n, coef, alpha};
    return b;
}

/* contracted (ij|kl) with optional extra powers (for derivative integrals) */
static double eri_bf_shift(const orc_basis *B, int i, int j, int k, int l, const int *di, const int *dj,
                           const int *dk, const int *dl, double (*wi)(double), double (*wj)(double),
                           double (*wk)(double), double (*wl)(double)) {
    double r = 0.0;
    for (int pi = 0; pi < B->nprim[i]; pi++) {
        int ii = B->istart[i] + pi;
        double ci = B->coef[ii] * (wi ? wi(B->alpha[ii]) : 1.0);
        for (int pj = 0; pj < B->nprim[j]; pj++) {
            int jj = B->istart[j] + pj;
            double cj = B->coef[jj] * (wj ? wj(B->alpha[jj]) : 1.0);
            for (int pk = 0; pk < B->nprim[k]; pk++) {
                int kk = B->istart[k] + pk;
                double ck = B->coef[kk] * (wk ? wk(B->alpha[kk]) : 1.0);
                for (int pl = 0; pl < B->nprim[l]; pl++) {
                    int ll = B->istart[l] + pl;
                    double cl = B->coef[ll] * (wl ? wl(B->alpha[ll]) : 1.0);
                    r += ci * cj * ck * cl *
                         orc_eri_prim(B->x[i], B->y[i], B->z[i], B->l[i] + di[0], B->m[i] + di[1], B->n[i] + di[2],
                                      B->alpha[ii], B->x[j], B->y[j], B->z[j], B->l[j] + dj[0], B->m[j] + dj[1],
                                      B->n[j] + dj[2], B->alpha[jj], B->x[k], B->y[k], B->z[k], B->l[k] + dk[0],
                                      B->m[k] + dk[1], B->n[k] + dk[2], B->alpha[kk], B->x[l], B->y[l], B->z[l],
                                      B->l[l] + dl[0], B->m[l] + dl[1], B->n[l] + dl[2], B->alpha[ll]);
                }
            }
        }
    }
    return r;
}

static const int Z3[3] = {0, 0, 0};

static double eri_bf(const orc_basis *B, int i, int j, int k, int l) {
    return eri_bf_shift(B, i, j, k, l, Z3, Z3, Z3, Z3, 0, 0, 0, 0);
}

double orc_eri_bf(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
                  const int *l, const int *m, const int *n, const double *coef, const double *alpha, int i, int j,
                  int k, int ll) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
    return eri_bf(&B, i, j, k, ll);
}

/* --------------------------------------------------- one-electron ------- */
void orc_getS(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
              const int *l, const int *m, const int *n, const double *coef, const double *alpha, double *S) {
    for (int j = 0; j < nbf; j++)
        for (int i = 0; i < nbf; i++) {
            double s = 0.0;
            for (int pj = 0; pj < nprim[j]; pj++) {
                int jj = istart[j] + pj;
                for (int pi = 0; pi < nprim[i]; pi++) {
                    int ii = istart[i] + pi;
                    s += coef[ii] * coef[jj] *
                         orc_overlap(alpha[ii], l[i], m[i], n[i], x[i], y[i], z[i], alpha[jj], l[j], m[j], n[j], x[j],
                                     y[j], z[j]);
                }
            }
            S[i + (size_t)j * nbf] = s;
        }
}

void orc_getT(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
              const int *l, const int *m, const int *n, const double *coef, const double *alpha, double *T) {
    for (int j = 0; j < nbf; j++)
        for (int i = 0; i < nbf; i++) {
            double s = 0.0;
            for (int pj = 0; pj < nprim[j]; pj++) {
                int jj = istart[j] + pj;
                for (int pi = 0; pi < nprim[i]; pi++) {
                    int ii = istart[i] + pi;
                    s += coef[ii] * coef[jj] *
                         orc_kinetic(alpha[ii], l[i], m[i], n[i], x[i], y[i], z[i], alpha[jj], l[j], m[j], n[j], x[j],
                                     y[j], z[j]);
                }
            }
            T[i + (size_t)j * nbf] = s;
        }
}

void orc_getV(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
              const int *l, const int *m, const int *n, const double *coef, const double *alpha, int nat,
              const int *atno, const double *ax, const double *ay, const double *az, double *V) {
    for (int j = 0; j < nbf; j++)
        for (int i = 0; i < nbf; i++) {
            double s = 0.0;
            for (int pj = 0; pj < nprim[j]; pj++) {
                int jj = istart[j] + pj;
                for (int pi = 0; pi < nprim[i]; pi++) {
                    int ii = istart[i] + pi;
                    for (int a = 0; a < nat; a++)
                        s += coef[ii] * coef[jj] * atno[a] *
                             orc_nuclear(x[i], y[i], z[i], l[i], m[i], n[i], alpha[ii], x[j], y[j], z[j], l[j], m[j],
                                         n[j], alpha[jj], ax[a], ay[a], az[a]);
                }
            }
            V[i + (size_t)j * nbf] = s;
        }
}

/* --------------------------------------------------- packed int2 -------- */
/* 0-based 8-fold packed index (cints.c:327-347), in 64-bit arithmetic. */
int64_t orc_ijkl2intindex(int64_t i, int64_t j, int64_t k, int64_t l) {
    int64_t t;
    if (i < j) { t = i; i = j; j = t; }
    if (k < l) { t = k; k = l; l = t; }
    int64_t ij = i * (i + 1) / 2 + j, kl = k * (k + 1) / 2 + l;
    if (ij < kl) { t = ij; ij = kl; kl = t; }
    return ij * (ij + 1) / 2 + kl;
}

void orc_getInts(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
                 const int *l, const int *m, const int *n, const double *coef, const double *alpha, double *Ints) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nbf; i++)
        for (int j = 0; j < i + 1; j++) {
            int64_t ij = (int64_t)i * (i + 1) / 2 + j;
            for (int k = 0; k < nbf; k++)
                for (int ll = 0; ll < k + 1; ll++) {
                    int64_t kl = (int64_t)k * (k + 1) / 2 + ll;
                    if (ij >= kl) Ints[orc_ijkl2intindex(i, j, k, ll)] = eri_bf(&B, i, j, k, ll);
                }
        }
}

/* ------------------------------------------------------- formG ---------- */
/* J_ij = sum_kl P_kl (ij|kl),  K_ij = sum_kl P_kl (il|kj)   (scf.f90:26-37,
 * TT(k,l,i,j) = (ij|kl) - 1/2 (il|kj)). Every unique (ij|kl) is evaluated
 * once and scattered to its 8 images. P, J, K: column-major n x n. */
void orc_jk(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
            const int *l, const int *m, const int *n, const double *coef, const double *alpha, const double *P,
            double *J, double *K, int nthreads) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
    size_t nn = (size_t)nbf * nbf;
    memset(J, 0, nn * sizeof(double));
    memset(K, 0, nn * sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *Jt = (double *)calloc(nn, sizeof(double));
        double *Kt = (double *)calloc(nn, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < nbf; i++)
            for (int j = 0; j <= i; j++)
                for (int k = 0; k <= i; k++)
                    for (int ll = 0; ll <= (k == i ? j : k); ll++) {
                        double v = eri_bf(&B, i, j, k, ll);
                        /* distinct ordered images of (ij|kl) */
                        int q[8][4] = {{i, j, k, ll}, {j, i, k, ll}, {i, j, ll, k}, {j, i, ll, k},
                                       {k, ll, i, j}, {ll, k, i, j}, {k, ll, j, i}, {ll, k, j, i}};
                        int nd = 0;
                        int seen[8][4];
                        for (int t = 0; t < 8; t++) {
                            int dup = 0;
                            for (int s = 0; s < nd; s++)
                                if (seen[s][0] == q[t][0] && seen[s][1] == q[t][1] && seen[s][2] == q[t][2] &&
                                    seen[s][3] == q[t][3])
                                    dup = 1;
                            if (dup) continue;
                            memcpy(seen[nd++], q[t], sizeof(q[t]));
                            int a = q[t][0], b = q[t][1], c = q[t][2], d = q[t][3];
                            /* (ab|cd): J_ab += P_cd v ; K_ad += P_cb v  (K_ij = sum P_kl (il|kj)) */
                            Jt[a + (size_t)b * nbf] += P[c + (size_t)d * nbf] * v;
                            Kt[a + (size_t)d * nbf] += P[c + (size_t)b * nbf] * v;
                        }
                    }
#pragma omp critical
        {
            for (size_t t = 0; t < nn; t++) {
                J[t] += Jt[t];
                K[t] += Kt[t];
            }
        }
        free(Jt);
        free(Kt);
    }
}

/* Scatters one unique integral v = (ij|kl) to J and K through its distinct
 * ordered images: (ab|cd) adds P_cd v to J_ab and P_cb v to K_ad. */
static void scatter_jk(int nbf, int i, int j, int k, int ll, double v, const double *P, double *J, double *K) {
    int q[8][4] = {{i, j, k, ll}, {j, i, k, ll}, {i, j, ll, k}, {j, i, ll, k},
                   {k, ll, i, j}, {ll, k, i, j}, {k, ll, j, i}, {ll, k, j, i}};
    int seen[8][4], nd = 0;
    for (int t = 0; t < 8; t++) {
        int dup = 0;
        for (int s = 0; s < nd; s++)
            if (seen[s][0] == q[t][0] && seen[s][1] == q[t][1] && seen[s][2] == q[t][2] && seen[s][3] == q[t][3])
                dup = 1;
        if (dup) continue;
        memcpy(seen[nd++], q[t], sizeof(q[t]));
        int a = q[t][0], b = q[t][1], c = q[t][2], d = q[t][3];
        J[a + (size_t)b * nbf] += P[c + (size_t)d * nbf] * v;
        K[a + (size_t)d * nbf] += P[c + (size_t)b * nbf] * v;
    }
}

/* Schwarz bounds of every shell pair (A >= B): Q_AB = sqrt(max |(ab|ab)|) over
 * the functions a of A and b of B, from the THO integrals. Shell s spans the
 * functions [first[s], first[s] + nf[s]). Q is packed: Q[A(A+1)/2 + B]. */
void orc_schwarz(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
                 const int *l, const int *m, const int *n, const double *coef, const double *alpha, int nsh,
                 const int *first, const int *nf, double *Q) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
#pragma omp parallel for schedule(dynamic, 1)
    for (int A = 0; A < nsh; A++)
        for (int Bs = 0; Bs <= A; Bs++) {
            double mx = 0.0;
            for (int a = first[A]; a < first[A] + nf[A]; a++)
                for (int b = first[Bs]; b < first[Bs] + nf[Bs]; b++) {
                    double v = fabs(eri_bf(&B, a, b, a, b));
                    if (v > mx) mx = v;
                }
            Q[(size_t)A * (A + 1) / 2 + Bs] = sqrt(mx);
        }
}

/* J and K of a density over Schwarz-screened shell quartets (the direct
 * engine's screening rule, no reference counterpart: the reference's formG,
 * scf.f90:26-37, contracts every stored integral). Unique shell quartets
 * (AB|CD) with A >= B, C >= D and pair index AB >= CD; a quartet is skipped
 * when Q_AB * Q_CD < thr (Q from orc_schwarz). Inside a quartet the unique
 * function quartets are those with i >= j when A == B, k >= l when C == D, and
 * function-pair index ij >= kl when AB == CD; each is evaluated once by THO
 * (same integrals as orc_jk) and scattered to its distinct images.
 * Returns the number of shell quartets evaluated. P, J, K column-major. */
double orc_jk_screened(int nbf, const int *nprim, const int *istart, const double *x, const double *y,
                       const double *z, const int *l, const int *m, const int *n, const double *coef,
                       const double *alpha, int nsh, const int *first, const int *nf, const double *Q, double thr,
                       const double *P, double *J, double *K, int nthreads) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
    size_t nn = (size_t)nbf * nbf;
    long npair = (long)nsh * (nsh + 1) / 2;
    memset(J, 0, nn * sizeof(double));
    memset(K, 0, nn * sizeof(double));
    double count = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : count)
    {
        double *Jt = (double *)calloc(nn, sizeof(double));
        double *Kt = (double *)calloc(nn, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (long ab = npair - 1; ab >= 0; ab--) {
            int A = (int)((sqrt(8.0 * ab + 1.0) - 1.0) / 2.0);
            while ((long)A * (A + 1) / 2 > ab) A--;
            while ((long)(A + 1) * (A + 2) / 2 <= ab) A++;
            int Bs = (int)(ab - (long)A * (A + 1) / 2);
            for (long cd = 0; cd <= ab; cd++) {
                if (Q[ab] * Q[cd] < thr) continue;
                int C = (int)((sqrt(8.0 * cd + 1.0) - 1.0) / 2.0);
                while ((long)C * (C + 1) / 2 > cd) C--;
                while ((long)(C + 1) * (C + 2) / 2 <= cd) C++;
                int D = (int)(cd - (long)C * (C + 1) / 2);
                count += 1.0;
                for (int i = first[A]; i < first[A] + nf[A]; i++)
                    for (int j = first[Bs]; j < first[Bs] + nf[Bs]; j++) {
                        if (A == Bs && j > i) continue;
                        long fij = (long)i * (i + 1) / 2 + j;
                        for (int k = first[C]; k < first[C] + nf[C]; k++)
                            for (int ll = first[D]; ll < first[D] + nf[D]; ll++) {
                                if (C == D && ll > k) continue;
                                if (ab == cd && (long)k * (k + 1) / 2 + ll > fij) continue;
                                scatter_jk(nbf, i, j, k, ll, eri_bf(&B, i, j, k, ll), P, Jt, Kt);
                            }
                    }
            }
        }
#pragma omp critical
        {
            for (size_t t = 0; t < nn; t++) {
                J[t] += Jt[t];
                K[t] += Kt[t];
            }
        }
        free(Jt);
        free(Kt);
    }
    return count;
}

/* ---------------------------------------------------- gradients --------- */
/* Primitive derivative weights for d/dA_x: 2*alpha * (l+1 term) - l * (l-1 term). */
static double w_two_alpha(double a) { return 2.0 * a; }

/* d(ij|kl)/dR for each of the 4 centres along axis ax: out[c] */
static void deri_bf(const orc_basis *B, int i, int j, int k, int l, int ax, double out[4]) {
    int up[3] = {0, 0, 0}, dn[3] = {0, 0, 0};
    up[ax] = 1;
    dn[ax] = -1;
    int pw[4] = {(ax == 0 ? B->l : ax == 1 ? B->m : B->n)[i], (ax == 0 ? B->l : ax == 1 ? B->m : B->n)[j],
                 (ax == 0 ? B->l : ax == 1 ? B->m : B->n)[k], (ax == 0 ? B->l : ax == 1 ? B->m : B->n)[l]};
    out[0] = eri_bf_shift(B, i, j, k, l, up, Z3, Z3, Z3, w_two_alpha, 0, 0, 0) -
             (pw[0] ? pw[0] * eri_bf_shift(B, i, j, k, l, dn, Z3, Z3, Z3, 0, 0, 0, 0) : 0.0);
    out[1] = eri_bf_shift(B, i, j, k, l, Z3, up, Z3, Z3, 0, w_two_alpha, 0, 0) -
             (pw[1] ? pw[1] * eri_bf_shift(B, i, j, k, l, Z3, dn, Z3, Z3, 0, 0, 0, 0) : 0.0);
    out[2] = eri_bf_shift(B, i, j, k, l, Z3, Z3, up, Z3, 0, 0, w_two_alpha, 0) -
             (pw[2] ? pw[2] * eri_bf_shift(B, i, j, k, l, Z3, Z3, dn, Z3, 0, 0, 0, 0) : 0.0);
    out[3] = eri_bf_shift(B, i, j, k, l, Z3, Z3, Z3, up, 0, 0, 0, w_two_alpha) -
             (pw[3] ? pw[3] * eri_bf_shift(B, i, j, k, l, Z3, Z3, Z3, dn, 0, 0, 0, 0) : 0.0);
}

/* Two-electron part of the RHF gradient:
 *   dE2/dX = sum_{ijkl} [1/2 P_ij P_kl - 1/4 P_ik P_jl] d(ij|kl)/dX
 * bf_atom[i] = atom index of basis function i; g: 3*nat, (x,y,z) per atom. */
void orc_grad_2e(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
                 const int *l, const int *m, const int *n, const double *coef, const double *alpha,
                 const int *bf_atom, int nat, const double *P, double *g, int nthreads) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
    memset(g, 0, 3 * (size_t)nat * sizeof(double));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *gt = (double *)calloc(3 * (size_t)nat, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < nbf; i++)
            for (int j = 0; j < nbf; j++)
                for (int k = 0; k < nbf; k++)
                    for (int ll = 0; ll < nbf; ll++) {
                        /* visit each unordered (ij|kl) through its canonical ordered form */
                        if (j > i || ll > k) continue;
                        int64_t ij = (int64_t)i * (i + 1) / 2 + j, kl = (int64_t)k * (k + 1) / 2 + ll;
                        if (kl > ij) continue;
                        double deg = (i == j ? 1.0 : 2.0) * (k == ll ? 1.0 : 2.0) * (ij == kl ? 1.0 : 2.0);
                        /* all ordered images carry the same derivative; sum gamma over them */
                        double gam =
                            deg * (0.5 * P[i + (size_t)j * nbf] * P[k + (size_t)ll * nbf] -
                                   0.125 * (P[i + (size_t)k * nbf] * P[j + (size_t)ll * nbf] +
                                            P[i + (size_t)ll * nbf] * P[j + (size_t)k * nbf]));
                        if (gam == 0.0) continue;
                        int cen[4] = {bf_atom[i], bf_atom[j], bf_atom[k], bf_atom[ll]};
                        for (int ax = 0; ax < 3; ax++) {
                            double d[4];
                            deri_bf(&B, i, j, k, ll, ax, d);
                            for (int c = 0; c < 4; c++) gt[3 * cen[c] + ax] += gam * d[c];
                        }
                    }
#pragma omp critical
        {
            for (int t = 0; t < 3 * nat; t++) g[t] += gt[t];
        }
        free(gt);
    }
}

/* One-electron part: sum_ij P_ij dH_ij/dX - sum_ij W_ij dS_ij/dX. dV includes
 * the operator (Hellmann-Feynman) term through translational invariance:
 * d/dC <a|1/r_C|b> = -(d/dA + d/dB) <a|1/r_C|b>. */
void orc_grad_1e(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
                 const int *l, const int *m, const int *n, const double *coef, const double *alpha,
                 const int *bf_atom, int nat, const int *atno, const double *ax_, const double *ay_,
                 const double *az_, const double *P, const double *W, double *g) {
    memset(g, 0, 3 * (size_t)nat * sizeof(double));
    double *dVc = (double *)calloc((size_t)nat, sizeof(double));
    for (int i = 0; i < nbf; i++)
        for (int j = 0; j < nbf; j++) {
            double pij = P[i + (size_t)j * nbf], wij = W[i + (size_t)j * nbf];
            for (int ax = 0; ax < 3; ax++) {
                /* derivative w.r.t. the centre of function i (function j's centre by symmetry of the sum) */
                double dS = 0, dT = 0, dVa = 0;
                for (int c = 0; c < nat; c++) dVc[c] = 0;
                for (int pi = 0; pi < nprim[i]; pi++) {
                    int ii = istart[i] + pi;
                    for (int pj = 0; pj < nprim[j]; pj++) {
                        int jj = istart[j] + pj;
                        double cc = coef[ii] * coef[jj], a = alpha[ii];
                        int li = l[i] + (ax == 0), mi = m[i] + (ax == 1), ni = n[i] + (ax == 2);
                        int lo = l[i] - (ax == 0), mo = m[i] - (ax == 1), no = n[i] - (ax == 2);
                        int pw = ax == 0 ? l[i] : ax == 1 ? m[i] : n[i];
                        dS += cc * (2 * a * orc_overlap(a, li, mi, ni, x[i], y[i], z[i], alpha[jj], l[j], m[j], n[j],
                                                        x[j], y[j], z[j]) -
                                    (pw ? pw * orc_overlap(a, lo, mo, no, x[i], y[i], z[i], alpha[jj], l[j], m[j],
                                                           n[j], x[j], y[j], z[j])
                                        : 0.0));
                        dT += cc * (2 * a * orc_kinetic(a, li, mi, ni, x[i], y[i], z[i], alpha[jj], l[j], m[j], n[j],
                                                        x[j], y[j], z[j]) -
                                    (pw ? pw * orc_kinetic(a, lo, mo, no, x[i], y[i], z[i], alpha[jj], l[j], m[j],
                                                           n[j], x[j], y[j], z[j])
                                        : 0.0));
                        for (int c = 0; c < nat; c++) {
                            double v = 2 * a * orc_nuclear(x[i], y[i], z[i], li, mi, ni, a, x[j], y[j], z[j], l[j],
                                                           m[j], n[j], alpha[jj], ax_[c], ay_[c], az_[c]) -
                                       (pw ? pw * orc_nuclear(x[i], y[i], z[i], lo, mo, no, a, x[j], y[j], z[j],
                                                              l[j], m[j], n[j], alpha[jj], ax_[c], ay_[c], az_[c])
                                           : 0.0);
                            v *= cc * atno[c];
                            dVa += v;
                            dVc[c] += v;
                        }
                    }
                }
                /* H, S, P, W symmetric: the derivative on function j's centre equals
                 * the (j,i) visit's derivative on its first centre, hence the 2. */
                g[3 * bf_atom[i] + ax] += 2.0 * (pij * (dT + dVa) - wij * dS);
                /* operator term d/dC = -(d/dA + d/dB), same factor 2 */
                for (int c = 0; c < nat; c++) g[3 * c + ax] -= 2.0 * pij * dVc[c];
            }
        }
    free(dVc);
}

/* ------------------------------------------------ CPU baseline sample --- */
/* Evaluates every contracted integral of nq shell quartets (quartet k spans
 * the basis functions [f[4k+x], f[4k+x] + nf[4k+x]) for x = 0..3) with the THO
 * primitives, and returns their sum (so nothing is optimised away). Used to
 * time the reference algorithm on a sample of a large build. */
double orc_eri_quartets(int nbf, const int *nprim, const int *istart, const double *x, const double *y,
                        const double *z, const int *l, const int *m, const int *n, const double *coef,
                        const double *alpha, int nq, const int *f, const int *nf, int nthreads) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
    double total = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int q = 0; q < nq; q++)
        for (int a = 0; a < nf[4 * q]; a++)
            for (int b = 0; b < nf[4 * q + 1]; b++)
                for (int c = 0; c < nf[4 * q + 2]; c++)
                    for (int d = 0; d < nf[4 * q + 3]; d++)
                        total += eri_bf(&B, f[4 * q] + a, f[4 * q + 1] + b, f[4 * q + 2] + c, f[4 * q + 3] + d);
    return total;
}

/* d(ij|kl)/dR_c along axis ax for the four centres c = 0..3 (functions i, j, k, l). */
void orc_deri_bf(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
                 const int *l, const int *m, const int *n, const double *coef, const double *alpha, int i, int j,
                 int k, int ll, int ax, double *out) {
    orc_basis B = mkbasis(nbf, nprim, istart, x, y, z, l, m, n, coef, alpha);
    deri_bf(&B, i, j, k, ll, ax, out);
}
