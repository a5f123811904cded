// Derivative electron-repulsion integrals contracted with the two-particle
// density, by Rys quadrature (host and device; no reference counterpart:
// hfsolver has no gradient code).
//
// For one shell quartet (AB|CD) and its slice (a-, c-components as in
// eri_quartet), accumulates
//   g_X += sum_{abcd} gam_abcd d(ab|cd)/dX,   X = A_x..A_z, B_x..B_z, C_x..C_z
// with the centre-derivative identity on every primitive,
//   d/dA_x x_A^i e^{-a x_A^2} = 2a x_A^{i+1} e^{..} - i x_A^{i-1} e^{..},
// so the x factor of d/dA_x (ab|cd) is 2 a I_x(i+1, ...) - i I_x(i-1, ...) with
// I_x the Rys 2-D integrals of eri_rys.h built one unit higher on the
// differentiated centre. The D derivative follows from translational
// invariance: d/dD = -(d/dA + d/dB + d/dC). For each root the bra centres A and
// B share one table built one power higher on the bra (a up to l_A + 1 and b
// up to l_B + 1 by the transfer), then C has its own; only one extended table
// is live at a time.
#pragma once

#include "eri_rys.h"

// Split policy of the gradient kernels: slice over a (then over c) while a
// thread would hold more than grad_split_limit density factors gam. The
// derivative tables need more registers than the energy's; a lower limit
// removes register spills but repeats the primitive set-up in every slice.
// 54 was the fastest of 81 / 54 / 36 on 64 waters over all classes; per class
// (launches serialised, 64 waters) FP64 (22|11) runs 42 % faster at 27 (its
// spills drop from 1212 to 116 bytes), FP32 (11|11) and (22|21) 27-28 %
// faster at 108 (no slice over a, resp. c); every other class is slower at
// 27 or at 108.
#ifndef HFX_GRAD_SPLIT
#define HFX_GRAD_SPLIT 54
#endif
HFX_HD constexpr int grad_split_limit(int la, int lb, int lc, int ld, bool fp32) {
    return fp32 ? (((la == 1 && lb == 1 && lc == 1 && ld == 1) || (la == 2 && lb == 2 && lc == 2 && ld == 1))
                       ? 2 * HFX_GRAD_SPLIT
                       : HFX_GRAD_SPLIT)
                : ((la == 2 && lb == 2 && lc == 1 && ld == 1) ? HFX_GRAD_SPLIT / 2 : HFX_GRAD_SPLIT);
}
HFX_HD constexpr bool gsplit_a(int la, int lb, int lc, int ld, bool fp32) {
    return ncart(la) > 1 && ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > grad_split_limit(la, lb, lc, ld, fp32);
}
HFX_HD constexpr bool gsplit_c(int la, int lb, int lc, int ld, bool fp32) {
    return gsplit_a(la, lb, lc, ld, fp32) && ncart(lc) > 1 &&
           ncart(lb) * ncart(lc) * ncart(ld) > grad_split_limit(la, lb, lc, ld, fp32);
}
HFX_HD constexpr int gnsplit(int la, int lb, int lc, int ld, bool fp32) {
    return (gsplit_a(la, lb, lc, ld, fp32) ? ncart(la) : 1) * (gsplit_c(la, lb, lc, ld, fp32) ? ncart(lc) : 1);
}

// Extents of the tables of the pass for centre CEN (0: A, 1: B, 2: C, 3: A and
// B from one table built one power higher on the bra).
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, int CEN>
struct GPass {
    static constexpr bool DA = CEN == 0 || CEN == 3, DB = CEN == 1 || CEN == 3;
    static constexpr int LAB = LA + LB + (DA || DB ? 1 : 0);  // highest power built on A
    static constexpr int LCD = LC + LD + (CEN == 2 ? 1 : 0); // highest power built on C
    static constexpr int LBx = LB + (DB ? 1 : 0);            // kept powers of B
    // kept powers of A: all 0..LA(+1), or the three around the selected one
    static constexpr int XA = SPA ? (DA ? 3 : 1) : (LA + 1 + (DA ? 1 : 0));
    static constexpr int XC = SPC ? (CEN == 2 ? 3 : 1) : (LC + 1 + (CEN == 2 ? 1 : 0));
};

// One direction of one root for the pass CEN: T[xa][ib][xc][id].
// With SPA (SPC) the kept a (c) powers are ea-1, ea, ea+1 for the
// differentiated centre (index 0, 1, 2; ea-1 < 0 reads 0 and is never used
// with a non-zero weight) and ea alone otherwise.
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, int CEN, typename R>
HFX_HD HFX_INLINE void grad_dir(R c00, R d00, R b10, R b01, R b00, R ab, R cd, R g00, int ea, int ec,
                                R (&T)[GPass<LA, LB, LC, LD, SPA, SPC, CEN>::XA][GPass<LA, LB, LC, LD, SPA, SPC, CEN>::LBx + 1]
                                      [GPass<LA, LB, LC, LD, SPA, SPC, CEN>::XC][LD + 1]) {
    using P = GPass<LA, LB, LC, LD, SPA, SPC, CEN>;
    constexpr int NAB = P::LAB, NCD = P::LCD, LBx = P::LBx;
    constexpr int LAk = LA + (P::DA ? 1 : 0);  // highest kept a power
    constexpr int LCk = LC + (CEN == 2 ? 1 : 0);
    R G[NAB + 1][NCD + 1];
    G[0][0] = g00;
    if (NAB > 0) G[1][0] = c00 * g00;
#pragma unroll
    for (int n = 1; n < NAB; n++) G[n + 1][0] = c00 * G[n][0] + R(n) * b10 * G[n - 1][0];
#pragma unroll
    for (int m = 0; m < NCD; m++) {
#pragma unroll
        for (int n = 0; n <= NAB; n++) {
            R v = d00 * G[n][m];
            if (m > 0) v += R(m) * b01 * G[n][m - 1];
            if (n > 0) v += R(n) * b00 * G[n - 1][m];
            G[n][m + 1] = v;
        }
    }
    R H[NAB + 1][LBx + 1][NCD + 1];
#pragma unroll
    for (int ia = 0; ia <= NAB; ia++)
#pragma unroll
        for (int m = 0; m <= NCD; m++) H[ia][0][m] = G[ia][m];
#pragma unroll
    for (int ib = 1; ib <= LBx; ib++)
#pragma unroll
        for (int ia = 0; ia <= NAB - ib; ia++)
#pragma unroll
            for (int m = 0; m <= NCD; m++) H[ia][ib][m] = H[ia + 1][ib - 1][m] + ab * H[ia][ib - 1][m];
    R HA[P::XA][LBx + 1][NCD + 1];
#pragma unroll
    for (int xa = 0; xa < P::XA; xa++)
#pragma unroll
        for (int ib = 0; ib <= LBx; ib++)
#pragma unroll
            for (int m = 0; m <= NCD; m++) {
                // entries with ia + ib > NAB are never built (and never used)
                if (SPA) {
                    const int off = P::DA ? xa - 1 : 0;
                    R v = R(0);
#pragma unroll
                    for (int t = 0; t <= LAk; t++)
                        if (t + ib <= NAB) v = (ea + off == t) ? H[t][ib][m] : v;
                    HA[xa][ib][m] = v;
                } else {
                    HA[xa][ib][m] = xa + ib <= NAB ? H[xa][ib][m] : R(0);
                }
            }
    R Kt[P::XA][LBx + 1][NCD + 1][LD + 1];
#pragma unroll
    for (int xa = 0; xa < P::XA; xa++)
#pragma unroll
        for (int ib = 0; ib <= LBx; ib++) {
#pragma unroll
            for (int ic = 0; ic <= NCD; ic++) Kt[xa][ib][ic][0] = HA[xa][ib][ic];
#pragma unroll
            for (int id = 1; id <= LD; id++)
#pragma unroll
                for (int ic = 0; ic <= NCD - id; ic++)
                    Kt[xa][ib][ic][id] = Kt[xa][ib][ic + 1][id - 1] + cd * Kt[xa][ib][ic][id - 1];
        }
#pragma unroll
    for (int xa = 0; xa < P::XA; xa++)
#pragma unroll
        for (int ib = 0; ib <= LBx; ib++)
#pragma unroll
            for (int xc = 0; xc < P::XC; xc++)
#pragma unroll
                for (int id = 0; id <= LD; id++) {
                    if (SPC) {
                        const int off = (CEN == 2) ? xc - 1 : 0;
                        R v = R(0);
#pragma unroll
                        for (int t = 0; t <= LCk; t++) v = (ec + off == t) ? Kt[xa][ib][t][id] : v;
                        T[xa][ib][xc][id] = v;
                    } else {
                        T[xa][ib][xc][id] = Kt[xa][ib][xc][id];
                    }
                }
}

// Contraction of one pass with gam for one root: adds to g[3*CEN + {0,1,2}].
// CEN = 3: both bra centres from one table, A into g[0..2] (exponent two_exp)
// and B into g[3..5] (exponent two_exp_b).
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, int CEN, typename R>
HFX_HD HFX_INLINE void grad_pass(const double* c00, const double* d00, R b10, R b01, R b00, const R* abv,
                                 const R* cdv, R g00z, const int* sel_a, const int* sel_c, R two_exp, R two_exp_b,
                                 const R (&gam)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT], R* g) {
    using P = GPass<LA, LB, LC, LD, SPA, SPC, CEN>;
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    R Tx[P::XA][P::LBx + 1][P::XC][LD + 1];
    R Ty[P::XA][P::LBx + 1][P::XC][LD + 1];
    R Tz[P::XA][P::LBx + 1][P::XC][LD + 1];
    grad_dir<LA, LB, LC, LD, SPA, SPC, CEN, R>(R(c00[0]), R(d00[0]), b10, b01, b00, abv[0], cdv[0], R(1), sel_a[0],
                                               sel_c[0], Tx);
    grad_dir<LA, LB, LC, LD, SPA, SPC, CEN, R>(R(c00[1]), R(d00[1]), b10, b01, b00, abv[1], cdv[1], R(1), sel_a[1],
                                               sel_c[1], Ty);
    grad_dir<LA, LB, LC, LD, SPA, SPC, CEN, R>(R(c00[2]), R(d00[2]), b10, b01, b00, abv[2], cdv[2], g00z, sel_a[2],
                                               sel_c[2], Tz);
    R gx = R(0), gy = R(0), gz = R(0);
    R hx = R(0), hy = R(0), hz = R(0);   // B components of CEN = 3
#pragma unroll
    for (int a = 0; a < Q::NAt; a++) {
        // exponents of a: table index of power p is p (unsliced) or 1 (+-1) around the selection
        const int ex = SPA ? sel_a[0] : cart_exp(LA, a, 0), ey = SPA ? sel_a[1] : cart_exp(LA, a, 1),
                  ez = SPA ? sel_a[2] : cart_exp(LA, a, 2);
        const int ax = SPA ? (P::DA ? 1 : 0) : cart_exp(LA, a, 0);
        const int ay = SPA ? (P::DA ? 1 : 0) : cart_exp(LA, a, 1);
        const int az = SPA ? (P::DA ? 1 : 0) : cart_exp(LA, a, 2);
#pragma unroll
        for (int b = 0; b < Q::NB; b++) {
            const int bx = cart_exp(LB, b, 0), by = cart_exp(LB, b, 1), bz = cart_exp(LB, b, 2);
#pragma unroll
            for (int c = 0; c < Q::NCt; c++) {
                const int fx = SPC ? sel_c[0] : cart_exp(LC, c, 0), fy = SPC ? sel_c[1] : cart_exp(LC, c, 1),
                          fz = SPC ? sel_c[2] : cart_exp(LC, c, 2);
                const int cx = SPC ? (CEN == 2 ? 1 : 0) : cart_exp(LC, c, 0);
                const int cy = SPC ? (CEN == 2 ? 1 : 0) : cart_exp(LC, c, 1);
                const int cz = SPC ? (CEN == 2 ? 1 : 0) : cart_exp(LC, c, 2);
#pragma unroll
                for (int d = 0; d < Q::ND; d++) {
                    const int dx = cart_exp(LD, d, 0), dy = cart_exp(LD, d, 1), dz = cart_exp(LD, d, 2);
                    const R gm = gam[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d];
                    const R x0 = Tx[ax][bx][cx][dx], y0 = Ty[ay][by][cy][dy], z0 = Tz[az][bz][cz][dz];
                    R dxv, dyv, dzv;
                    if (CEN == 3) {
                        const R ux = two_exp_b * Tx[ax][bx + 1][cx][dx] - (bx > 0 ? R(bx) * Tx[ax][bx > 0 ? bx - 1 : 0][cx][dx] : R(0));
                        const R uy = two_exp_b * Ty[ay][by + 1][cy][dy] - (by > 0 ? R(by) * Ty[ay][by > 0 ? by - 1 : 0][cy][dy] : R(0));
                        const R uz = two_exp_b * Tz[az][bz + 1][cz][dz] - (bz > 0 ? R(bz) * Tz[az][bz > 0 ? bz - 1 : 0][cz][dz] : R(0));
                        hx += gm * ux * y0 * z0;
                        hy += gm * x0 * uy * z0;
                        hz += gm * x0 * y0 * uz;
                    }
                    if (CEN == 0 || CEN == 3) {
                        dxv = two_exp * Tx[ax + 1][bx][cx][dx] - (ex > 0 ? R(ex) * Tx[ax > 0 ? ax - 1 : 0][bx][cx][dx] : R(0));
                        dyv = two_exp * Ty[ay + 1][by][cy][dy] - (ey > 0 ? R(ey) * Ty[ay > 0 ? ay - 1 : 0][by][cy][dy] : R(0));
                        dzv = two_exp * Tz[az + 1][bz][cz][dz] - (ez > 0 ? R(ez) * Tz[az > 0 ? az - 1 : 0][bz][cz][dz] : R(0));
                    } else if (CEN == 1) {
                        dxv = two_exp * Tx[ax][bx + 1][cx][dx] - (bx > 0 ? R(bx) * Tx[ax][bx > 0 ? bx - 1 : 0][cx][dx] : R(0));
                        dyv = two_exp * Ty[ay][by + 1][cy][dy] - (by > 0 ? R(by) * Ty[ay][by > 0 ? by - 1 : 0][cy][dy] : R(0));
                        dzv = two_exp * Tz[az][bz + 1][cz][dz] - (bz > 0 ? R(bz) * Tz[az][bz > 0 ? bz - 1 : 0][cz][dz] : R(0));
                    } else {
                        dxv = two_exp * Tx[ax][bx][cx + 1][dx] - (fx > 0 ? R(fx) * Tx[ax][bx][cx > 0 ? cx - 1 : 0][dx] : R(0));
                        dyv = two_exp * Ty[ay][by][cy + 1][dy] - (fy > 0 ? R(fy) * Ty[ay][by][cy > 0 ? cy - 1 : 0][dy] : R(0));
                        dzv = two_exp * Tz[az][bz][cz + 1][dz] - (fz > 0 ? R(fz) * Tz[az][bz][cz > 0 ? cz - 1 : 0][dz] : R(0));
                    }
                    gx += gm * dxv * y0 * z0;
                    gy += gm * x0 * dyv * z0;
                    gz += gm * x0 * y0 * dzv;
                }
            }
        }
    }
    const int g0 = CEN == 3 ? 0 : 3 * CEN;
    g[g0 + 0] += gx;
    g[g0 + 1] += gy;
    g[g0 + 2] += gz;
    if (CEN == 3) {
        g[3] += hx;
        g[4] += hy;
        g[5] += hz;
    }
}

// g[0..8] += sum_{abcd in slice} gam_abcd d(ab|cd)/d(A, B, C) for the quartet.
// bra: shell pair (primitive pairs prims[bra.p0 ..], exponents pexp[bra.p0 ..]);
// ket: C - D = cd[], primitive pairs and C exponents from kp (outer loop, one
// per-lane load each; the bra's are the same for the whole wavefront).
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R, class KP>
HFX_HD HFX_INLINE void grad_quartet_k(const ShellPair& bra, const PrimPair* __restrict__ prims,
                                      const PrimExp* __restrict__ pexp, double cdx, double cdy, double cdz,
                                      const KP& kp, int sa, int sc, const RysTables& tab,
                                      const R (&gam)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT], double (&g)[9]) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
    int sel_a[3] = {0, 0, 0}, sel_c[3] = {0, 0, 0};
    if (SPA) {
#pragma unroll
        for (int i = 0; i < Q::NA; i++)
            if (sa == i) {
                sel_a[0] = cart_exp(LA, i, 0);
                sel_a[1] = cart_exp(LA, i, 1);
                sel_a[2] = cart_exp(LA, i, 2);
            }
    }
    if (SPC) {
#pragma unroll
        for (int i = 0; i < Q::NC; i++)
            if (sc == i) {
                sel_c[0] = cart_exp(LC, i, 0);
                sel_c[1] = cart_exp(LC, i, 1);
                sel_c[2] = cart_exp(LC, i, 2);
            }
    }
    const R abv[3] = {R(bra.ABx), R(bra.ABy), R(bra.ABz)};
    const R cdv[3] = {R(cdx), R(cdy), R(cdz)};
    R gl[9];
#pragma unroll
    for (int k = 0; k < 9; k++) gl[k] = R(0);
    const int nkp = kp.count();
    for (int k1 = 0; k1 < nkp; k1++) {
        const auto Qp = kp.get(k1);
        const R two_c = R(2.0 * double(kp.cexp(k1)));
        for (int i1 = 0; i1 < bra.np; i1++) {
            const PrimPair P = hfx_ld_prim(prims + bra.p0 + i1);
            const PrimExp EB = hfx_ld_pexp(pexp + bra.p0 + i1);
            const double p = P.p, q = double(Qp.p), pq = p + q;
            const double ip = P.ip, iq = double(Qp.ip);
            const double isq = hfx_rsqrt(pq);
            const double ipq = isq * isq;
            const double rho = p * q * ipq;
            const double PQ[3] = {P.Px - double(Qp.Px), P.Py - double(Qp.Py), P.Pz - double(Qp.Pz)};
            const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            const double pref = HFX_TWO_PI_52 * P.K * double(Qp.K) * ip * iq * isq;
            double u[NR], w[NR];
            if constexpr (sizeof(R) == 4 && NR <= 2) {
                rys_roots_small<NR, HFX_BOYS_K32>(T, u, w, tab.boys);
            } else if constexpr (sizeof(R) == 4 && HFX_FP32_RULES >= 1) {
                // single-precision derivative integrals: the rules of the FP32 J/K kernels
                float uf[NR], wf[NR];
                rys_roots_f<NR>(float(T), uf, wf, tab);
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    u[r] = uf[r];
                    w[r] = wf[r];
                }
            } else if constexpr (NR <= 2) {
                rys_roots_small<NR>(T, u, w, tab.boys);
            } else {
                rys_roots<NR>(T, u, w, tab);
            }
            const double PA[3] = {P.PAx, P.PAy, P.PAz};
            const double QC[3] = {double(Qp.QCx), double(Qp.QCy), double(Qp.QCz)};
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const double ur = u[r];
                const double t1 = ur * rho;
                const R b00 = R(0.5 * ur * ipq);
                const R b10 = R(0.5 * ip * (1.0 - t1 * ip));
                const R b01 = R(0.5 * iq * (1.0 - t1 * iq));
                const double fp = t1 * ip, fq = t1 * iq;
                const double c00[3] = {PA[0] - fp * PQ[0], PA[1] - fp * PQ[1], PA[2] - fp * PQ[2]};
                const double d00[3] = {QC[0] + fq * PQ[0], QC[1] + fq * PQ[1], QC[2] + fq * PQ[2]};
                const R wz = R(w[r] * pref);
                grad_pass<LA, LB, LC, LD, SPA, SPC, 3, R>(c00, d00, b10, b01, b00, abv, cdv, wz, sel_a, sel_c,
                                                          R(2.0 * EB.aa), R(2.0 * EB.ab), gam, gl);
                grad_pass<LA, LB, LC, LD, SPA, SPC, 2, R>(c00, d00, b10, b01, b00, abv, cdv, wz, sel_a, sel_c,
                                                          two_c, R(0), gam, gl);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) g[k] += double(gl[k]);
}

// The same for a ket given as a shell pair with PrimPair / PrimExp records.
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void grad_quartet(const ShellPair& bra, const ShellPair& ket, const PrimPair* __restrict__ prims,
                                    const PrimExp* __restrict__ pexp, int sa, int sc, const RysTables& tab,
                                    const R (&gam)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT], double (&g)[9]) {
    KetPrimsAoSG kp;
    kp.pp = prims + ket.p0;
    kp.np = ket.np;
    kp.pe = pexp + ket.p0;
    grad_quartet_k<LA, LB, LC, LD, SPA, SPC, R>(bra, prims, pexp, ket.ABx, ket.ABy, ket.ABz, kp, sa, sc, tab, gam, g);
}

// gam_abcd = s (4 D_ab D_cd - D_ac D_bd - D_ad D_bc) for the thread's slice;
// fa / fc are the first functions of the slice (fa + sa, fc + sc when slicing).
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void grad_gamma(int fa, int fb, int fc, int fd, const double* __restrict__ D, int n, double s,
                                  R (&gam)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT]) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
#pragma unroll
    for (int a = 0; a < Q::NAt; a++)
#pragma unroll
        for (int b = 0; b < Q::NB; b++)
#pragma unroll
            for (int c = 0; c < Q::NCt; c++)
#pragma unroll
                for (int d = 0; d < Q::ND; d++) {
                    const double dab = D[(size_t)(fa + a) * n + fb + b], dcd = D[(size_t)(fc + c) * n + fd + d];
                    const double dac = D[(size_t)(fa + a) * n + fc + c], dbd = D[(size_t)(fb + b) * n + fd + d];
                    const double dad = D[(size_t)(fa + a) * n + fd + d], dbc = D[(size_t)(fb + b) * n + fc + c];
                    gam[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d] = R(s * (4.0 * dab * dcd - dac * dbd - dad * dbc));
                }
}
