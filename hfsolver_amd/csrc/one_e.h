// One-electron integrals of a primitive pair -- overlap, kinetic energy,
// nuclear attraction -- and their derivatives with respect to the centre of
// the first function, written once for host and device (templates on the
// angular momenta).
//
// Counterparts of hfsolver's getS / getT / getV (src/qc.c:6-93), which
// evaluate overlap (src/cints.c:212-244), kinetic (:186-210) and nuclear
// attraction (:246-307) one basis-function pair at a time by the THO closed
// forms. Here:
//   overlap     1-D Obara-Saika tables S(i, j) per Cartesian direction,
//               S(i+1, j) = PA S(i, j) + (i S(i-1, j) + j S(i, j-1)) / 2p;
//   kinetic     -1/2 d^2/dx^2 on the second function:
//               T_x(i, j) = -2 b^2 S(i, j+2) + b (2j+1) S(i, j) - j(j-1)/2 S(i, j-2);
//   nuclear     Rys quadrature with n = (i + j)/2 + 1 roots of F(p |P - C|^2),
//               2-D table G(k) = C00 G(k-1) + (k-1) B10 G(k-2), C00 = PA - u PC,
//               B10 = (1 - u)/(2p), transferred to (i, j) with A - B;
//   derivative  d/dA_x x_A^i e^{-a x_A^2} = 2a x_A^{i+1} e^{..} - i x_A^{i-1} e^{..}
//               (tables built one power higher on A); the Hellmann-Feynman
//               term of nucleus C is -(d/dA + d/dB) of the same integrals.
#pragma once

#include "rys.h"

// 1-D overlap table S[i][j], i <= IM, j <= JM, without the (pi/p)^{1/2} factor.
template <int IM, int JM>
HFX_HD HFX_INLINE void oe_overlap_1d(double PA, double PB, double oo2p, double (&S)[IM + 1][JM + 1]) {
#pragma unroll
    for (int i = 0; i <= IM; i++)
#pragma unroll
        for (int j = 0; j <= JM; j++) {
            double v;
            if (i == 0 && j == 0) {
                v = 1.0;
            } else if (i > 0) {
                v = PA * S[i - 1][j];
                if (i > 1) v += (i - 1) * oo2p * S[i - 2][j];
                if (j > 0) v += j * oo2p * S[i - 1][j - 1];
            } else {
                v = PB * S[i][j - 1];
                if (j > 1) v += (j - 1) * oo2p * S[i][j - 2];
            }
            S[i][j] = v;
        }
}

// Nuclear-attraction 2-D table of one root: I[i][j], i <= IM, j <= JM.
template <int IM, int JM>
HFX_HD HFX_INLINE void oe_nuc_1d(double c00, double b10, double AB, double g0, double (&I)[IM + 1][JM + 1]) {
    constexpr int NM = IM + JM;
    double G[NM + 1];
    G[0] = g0;
    if (NM > 0) G[1] = c00 * g0;
#pragma unroll
    for (int k = 1; k < NM; k++) G[k + 1] = c00 * G[k] + k * b10 * G[k - 1];
    double H[NM + 1][JM + 1];
#pragma unroll
    for (int k = 0; k <= NM; k++) H[k][0] = G[k];
#pragma unroll
    for (int j = 1; j <= JM; j++)
#pragma unroll
        for (int k = 0; k <= NM - j; k++) H[k][j] = H[k + 1][j - 1] + AB * H[k][j - 1];
#pragma unroll
    for (int i = 0; i <= IM; i++)
#pragma unroll
        for (int j = 0; j <= JM; j++) I[i][j] = H[i][j];
}

// Geometry of one primitive pair.
struct OePrim {
    double a, b, p, oo2p, K;  // exponents, a + b, 1/2p, c_a c_b e^{-ab/p |A-B|^2}
    double PA[3], PB[3], P[3];
};

// 1-D overlap tables of one primitive pair, I + 1 powers on A, J + 1 on B.
template <int IM, int JM>
HFX_HD HFX_INLINE void oe_tables(const OePrim& q, double (&S1)[3][IM + 1][JM + 1]) {
#pragma unroll
    for (int d = 0; d < 3; d++) oe_overlap_1d<IM, JM>(q.PA[d], q.PB[d], q.oo2p, S1[d]);
}

// 1-D kinetic factor -1/2 d^2/dx^2 acting on the second function.
template <int IM, int JM>
HFX_HD HFX_INLINE double oe_tk(const double (&S1)[3][IM + 1][JM + 1], double b, int d, int i, int j) {
    double v = -2.0 * b * b * S1[d][i][j + 2] + b * (2 * j + 1) * S1[d][i][j];
    if (j >= 2) v -= 0.5 * j * (j - 1) * S1[d][i][j - 2];
    return v;
}

// Overlap and kinetic contributions of one primitive pair: S[ia][ib], T[ia][ib]
// (function order of the Cartesian shells).
template <int LA, int LB>
HFX_HD HFX_INLINE void oe_st(const OePrim& q, double (&S)[ncart(LA)][ncart(LB)], double (&T)[ncart(LA)][ncart(LB)]) {
    constexpr int IM = LA, JM = LB + 2;
    double S1[3][IM + 1][JM + 1];
    oe_tables<IM, JM>(q, S1);
    const double s0 = HFX_PI / q.p * sqrt(HFX_PI / q.p) * q.K;
#pragma unroll
    for (int ia = 0; ia < ncart(LA); ia++) {
        const int ax = cart_exp(LA, ia, 0), ay = cart_exp(LA, ia, 1), az = cart_exp(LA, ia, 2);
#pragma unroll
        for (int ib = 0; ib < ncart(LB); ib++) {
            const int bx = cart_exp(LB, ib, 0), by = cart_exp(LB, ib, 1), bz = cart_exp(LB, ib, 2);
            const double sx = S1[0][ax][bx], sy = S1[1][ay][by], sz = S1[2][az][bz];
            S[ia][ib] += s0 * sx * sy * sz;
            T[ia][ib] += s0 * (oe_tk<IM, JM>(S1, q.b, 0, ax, bx) * sy * sz + sx * oe_tk<IM, JM>(S1, q.b, 1, ay, by) * sz +
                               sx * sy * oe_tk<IM, JM>(S1, q.b, 2, az, bz));
        }
    }
}

// Gradient form: g[d] += sum_ab (pw_ab dT_ab/dA_d - ww_ab dS_ab/dA_d).
template <int LA, int LB>
HFX_HD HFX_INLINE void oe_st_grad(const OePrim& q, const double (&pw)[ncart(LA)][ncart(LB)],
                                  const double (&ww)[ncart(LA)][ncart(LB)], double (&g)[3]) {
    constexpr int IM = LA + 1, JM = LB + 2;
    double S1[3][IM + 1][JM + 1];
    oe_tables<IM, JM>(q, S1);
    const double s0 = HFX_PI / q.p * sqrt(HFX_PI / q.p) * q.K;
    // s and t of the product x^ex y^ey z^ez (A side) with (bx, by, bz)
    auto st = [&](int ex, int ey, int ez, int bx, int by, int bz, double& s, double& t) {
        const double sx = S1[0][ex][bx], sy = S1[1][ey][by], sz = S1[2][ez][bz];
        s = s0 * sx * sy * sz;
        t = s0 * (oe_tk<IM, JM>(S1, q.b, 0, ex, bx) * sy * sz + sx * oe_tk<IM, JM>(S1, q.b, 1, ey, by) * sz +
                  sx * sy * oe_tk<IM, JM>(S1, q.b, 2, ez, bz));
    };
#pragma unroll
    for (int ia = 0; ia < ncart(LA); ia++) {
        const int e[3] = {cart_exp(LA, ia, 0), cart_exp(LA, ia, 1), cart_exp(LA, ia, 2)};
#pragma unroll
        for (int ib = 0; ib < ncart(LB); ib++) {
            const int bx = cart_exp(LB, ib, 0), by = cart_exp(LB, ib, 1), bz = cart_exp(LB, ib, 2);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double su, tu, sd = 0.0, td = 0.0;
                st(e[0] + (d == 0), e[1] + (d == 1), e[2] + (d == 2), bx, by, bz, su, tu);
                if (e[d] > 0) st(e[0] - (d == 0), e[1] - (d == 1), e[2] - (d == 2), bx, by, bz, sd, td);
                const double ds = 2 * q.a * su - e[d] * sd, dt = 2 * q.a * tu - e[d] * td;
                g[d] += pw[ia][ib] * dt - ww[ia][ib] * ds;
            }
        }
    }
}

// 2-D tables of the nuclear attraction by the charge Z at C, per root.
template <int IM, int JM, int NR>
HFX_HD HFX_INLINE void oe_v_tables(const OePrim& q, const double* C, double Z, const double* AB, const RysTables& tab,
                                   double (&I)[NR][3][IM + 1][JM + 1]) {
    const double PC[3] = {q.P[0] - C[0], q.P[1] - C[1], q.P[2] - C[2]};
    const double T = q.p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
    double u[NR], w[NR];
    if constexpr (NR <= 2)
        rys_roots_small<NR>(T, u, w, tab.boys);
    else
        rys_roots<NR>(T, u, w, tab);
    const double pref = -Z * 2.0 * HFX_PI / q.p * q.K;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const double b10 = (1.0 - u[r]) * q.oo2p;
#pragma unroll
        for (int d = 0; d < 3; d++)
            oe_nuc_1d<IM, JM>(q.PA[d] - u[r] * PC[d], b10, AB[d], d == 2 ? w[r] * pref : 1.0, I[r][d]);
    }
}

// Nuclear attraction of one primitive pair by the charge Z at C:
// V[ia][ib] += -Z <a|1/r_C|b>.
template <int LA, int LB>
HFX_HD HFX_INLINE void oe_v(const OePrim& q, const double* C, double Z, const double* AB, const RysTables& tab,
                            double (&V)[ncart(LA)][ncart(LB)]) {
    constexpr int IM = LA, JM = LB, NR = (IM + JM) / 2 + 1;
    double I[NR][3][IM + 1][JM + 1];
    oe_v_tables<IM, JM, NR>(q, C, Z, AB, tab, I);
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
        for (int ia = 0; ia < ncart(LA); ia++)
#pragma unroll
            for (int ib = 0; ib < ncart(LB); ib++)
                V[ia][ib] += I[r][0][cart_exp(LA, ia, 0)][cart_exp(LB, ib, 0)] *
                             I[r][1][cart_exp(LA, ia, 1)][cart_exp(LB, ib, 1)] *
                             I[r][2][cart_exp(LA, ia, 2)][cart_exp(LB, ib, 2)];
}

// Gradient form: h[d] += sum_ab pw_ab d/dA_d (-Z <a|1/r_C|b>).
template <int LA, int LB>
HFX_HD HFX_INLINE void oe_v_grad(const OePrim& q, const double* C, double Z, const double* AB, const RysTables& tab,
                                 const double (&pw)[ncart(LA)][ncart(LB)], double (&h)[3]) {
    constexpr int IM = LA + 1, JM = LB, NR = (IM + JM) / 2 + 1;
    double I[NR][3][IM + 1][JM + 1];
    oe_v_tables<IM, JM, NR>(q, C, Z, AB, tab, I);
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
        for (int ia = 0; ia < ncart(LA); ia++) {
            const int e[3] = {cart_exp(LA, ia, 0), cart_exp(LA, ia, 1), cart_exp(LA, ia, 2)};
#pragma unroll
            for (int ib = 0; ib < ncart(LB); ib++) {
                const int b3[3] = {cart_exp(LB, ib, 0), cart_exp(LB, ib, 1), cart_exp(LB, ib, 2)};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double up = 2 * q.a, dn = 0.0;
#pragma unroll
                    for (int x = 0; x < 3; x++) up *= I[r][x][e[x] + (x == d)][b3[x]];
                    if (e[d] > 0) {
                        dn = e[d];
#pragma unroll
                        for (int x = 0; x < 3; x++) dn *= I[r][x][x == d ? e[x] - 1 : e[x]][b3[x]];
                    }
                    h[d] += pw[ia][ib] * (up - dn);
                }
            }
        }
}
