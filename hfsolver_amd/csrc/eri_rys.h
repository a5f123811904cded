// Contracted shell-quartet electron-repulsion integrals by Rys quadrature,
// written once for host and device (template on the angular momenta and on
// the arithmetic type R = double or float).
//
// Counterpart of hfsolver's contr_coulomb / coulomb_repulsion
// (src/cints.c:72-149), which evaluate the same integrals one basis-function
// quartet at a time by the THO closed form. Here one call evaluates a whole
// (or a slice of a) Cartesian shell quartet (AB|CD), l_A >= l_B, l_C >= l_D:
//
//   (ab|cd) = sum_{prim pairs} 2 pi^{5/2} K_ab K_cd / (p q sqrt(p+q))
//             * sum_{roots r} w_r Ix(r) Iy(r) Iz(r)
//
// with the 2-D integrals built by the Rys recurrences
//   G(n+1,0) = C00 G(n,0) + n B10 G(n-1,0)
//   G(n,m+1) = C00' G(n,m) + m B01 G(n,m-1) + n B00 G(n-1,m)
//   C00 = PA - u rho/p PQ,  C00' = QC + u rho/q PQ,
//   B10 = (1 - u rho/p)/(2p),  B01 = (1 - u rho/q)/(2q),  B00 = u/(2(p+q)),
// followed by the horizontal transfers I(a,b+1) = I(a+1,b) + AB I(a,b)
// (bra) and the same on the ket with CD.
//
// Splitting: for the large classes one thread evaluates only the slice of a
// single a-component (SPA) and/or a single c-component (SPC); the component is
// a run-time, wavefront-uniform index, selected with conditional moves so that
// every array stays in registers.
#pragma once

#include "hfx_common.h"
#include "rys.h"

template <int LA, int LB, int LC, int LD, bool SPA, bool SPC>
struct QClass {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NAt = SPA ? 1 : NA;
    static constexpr int NCt = SPC ? 1 : NC;
    static constexpr int NOUT = NAt * NB * NCt * ND;
    static constexpr int LAB = LA + LB, LCD = LC + LD;
    static constexpr int NR = (LAB + LCD) / 2 + 1;
    static constexpr int XA = SPA ? 1 : LA + 1;   // a-exponent extent kept after selection
    static constexpr int XC = SPC ? 1 : LC + 1;
};

// Split policy of the engine: slice over a (then over c) while a thread would
// hold more integrals than the limit of its arithmetic type. Every slice
// repeats the primitive set-up, the Rys rule and the recurrences, so fewer
// slices pay as long as the accumulators fit in registers: FP64 108 (64
// waters, FP64 build: 0.1352 s at 81, 0.1289 s at 108, 0.1373 s at 162;
// 54 and 27 slower than 81), FP32 162 (registers of half the width).
#ifndef HFX_JK_SPLIT_F64
#define HFX_JK_SPLIT_F64 108
#endif
#ifndef HFX_JK_SPLIT_F32
#define HFX_JK_SPLIT_F32 162
#endif
HFX_HD constexpr int jk_split_limit(bool fp32) { return fp32 ? HFX_JK_SPLIT_F32 : HFX_JK_SPLIT_F64; }
// Classes sliced over their c-components alone although they fit the limit:
// FP64 (22|10) (108 integrals per thread spill 744 bytes; three slices of 36).
HFX_HD constexpr bool split_c_only(int la, int lb, int lc, int ld, bool fp32) {
    return !fp32 && la == 2 && lb == 2 && lc == 1 && ld == 0;
}
HFX_HD constexpr bool split_a(int la, int lb, int lc, int ld, bool fp32 = false) {
    return !split_c_only(la, lb, lc, ld, fp32) && ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > jk_split_limit(fp32);
}
HFX_HD constexpr bool split_c(int la, int lb, int lc, int ld, bool fp32 = false) {
    return split_c_only(la, lb, lc, ld, fp32) ||
           (split_a(la, lb, lc, ld, fp32) && ncart(lb) * ncart(lc) * ncart(ld) > jk_split_limit(fp32));
}
HFX_HD constexpr int nsplit(int la, int lb, int lc, int ld, bool fp32 = false) {
    return (split_a(la, lb, lc, ld, fp32) ? ncart(la) : 1) * (split_c(la, lb, lc, ld, fp32) ? ncart(lc) : 1);
}
// Narrow row launches, for molecules whose K rows outgrow the LDS of a block:
// a class with several a-components that the policy above leaves unsliced over
// a is sliced over them (then over c by the same limit), so that a row block
// holds one K row of the a-shell instead of ncart(la).
HFX_HD constexpr bool narrow_differs(int la, int lb, int lc, int ld, bool fp32) {
    return ncart(la) > 1 && !split_a(la, lb, lc, ld, fp32);
}
HFX_HD constexpr bool split_a_rows(int la, int lb, int lc, int ld, bool fp32, bool narrow) {
    return (narrow && narrow_differs(la, lb, lc, ld, fp32)) || split_a(la, lb, lc, ld, fp32);
}
HFX_HD constexpr bool split_c_rows(int la, int lb, int lc, int ld, bool fp32, bool narrow) {
    return narrow && narrow_differs(la, lb, lc, ld, fp32) ? ncart(lb) * ncart(lc) * ncart(ld) > jk_split_limit(fp32)
                                                          : split_c(la, lb, lc, ld, fp32);
}
HFX_HD constexpr int nsplit_rows(int la, int lb, int lc, int ld, bool fp32, bool narrow) {
    return (split_a_rows(la, lb, lc, ld, fp32, narrow) ? ncart(la) : 1) *
           (split_c_rows(la, lb, lc, ld, fp32, narrow) ? ncart(lc) : 1);
}

// One Cartesian direction of one Rys root: 2-D integrals, bra and ket
// transfers, selection of the a / c exponent (ea, ec) when slicing.
// out[ia][ib][ic][id] for ia < XA, ib <= LB, ic < XC, id <= LD.
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void rys_dir(R c00, R d00, R b10, R b01, R b00, R ab, R cd, R g00, int ea, int ec,
                               R (&out)[QClass<LA, LB, LC, LD, SPA, SPC>::XA][LB + 1]
                                       [QClass<LA, LB, LC, LD, SPA, SPC>::XC][LD + 1]) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int LAB = Q::LAB, LCD = Q::LCD;
    R G[LAB + 1][LCD + 1];
    G[0][0] = g00;
    if constexpr (LAB > 0) G[1][0] = c00 * g00;
#pragma unroll
    for (int n = 1; n < LAB; n++) G[n + 1][0] = c00 * G[n][0] + R(n) * b10 * G[n - 1][0];
#pragma unroll
    for (int m = 0; m < LCD; m++) {
#pragma unroll
        for (int n = 0; n <= LAB; n++) {
            R v = d00 * G[n][m];
            if (m > 0) v += R(m) * b01 * G[n][m - 1];
            if (n > 0) v += R(n) * b00 * G[n - 1][m];
            G[n][m + 1] = v;
        }
    }
    // bra transfer: H[ia][ib][m], ia + ib <= LAB
    R H[LAB + 1][LB + 1][LCD + 1];
#pragma unroll
    for (int ia = 0; ia <= LAB; ia++)
#pragma unroll
        for (int m = 0; m <= LCD; m++) H[ia][0][m] = G[ia][m];
#pragma unroll
    for (int ib = 1; ib <= LB; ib++)
#pragma unroll
        for (int ia = 0; ia <= LAB - ib; ia++)
#pragma unroll
            for (int m = 0; m <= LCD; m++) H[ia][ib][m] = H[ia + 1][ib - 1][m] + ab * H[ia][ib - 1][m];
    // keep ia <= LA (or the selected ia = ea)
    R HA[Q::XA][LB + 1][LCD + 1];
#pragma unroll
    for (int xa = 0; xa < Q::XA; xa++)
#pragma unroll
        for (int ib = 0; ib <= LB; ib++)
#pragma unroll
            for (int m = 0; m <= LCD; m++) {
                if (SPA) {
                    R v = H[0][ib][m];
#pragma unroll
                    for (int t = 1; t <= LA; t++) v = (ea == t) ? H[t][ib][m] : v;
                    HA[xa][ib][m] = v;
                } else {
                    HA[xa][ib][m] = H[xa][ib][m];
                }
            }
    // ket transfer: K[xa][ib][ic][id], ic + id <= LCD
    R Kt[Q::XA][LB + 1][LCD + 1][LD + 1];
#pragma unroll
    for (int xa = 0; xa < Q::XA; xa++)
#pragma unroll
        for (int ib = 0; ib <= LB; ib++) {
#pragma unroll
            for (int ic = 0; ic <= LCD; ic++) Kt[xa][ib][ic][0] = HA[xa][ib][ic];
#pragma unroll
            for (int id = 1; id <= LD; id++)
#pragma unroll
                for (int ic = 0; ic <= LCD - id; ic++)
                    Kt[xa][ib][ic][id] = Kt[xa][ib][ic + 1][id - 1] + cd * Kt[xa][ib][ic][id - 1];
        }
#pragma unroll
    for (int xa = 0; xa < Q::XA; xa++)
#pragma unroll
        for (int ib = 0; ib <= LB; ib++)
#pragma unroll
            for (int xc = 0; xc < Q::XC; xc++)
#pragma unroll
                for (int id = 0; id <= LD; id++) {
                    if (SPC) {
                        R v = Kt[xa][ib][0][id];
#pragma unroll
                        for (int t = 1; t <= LC; t++) v = (ec == t) ? Kt[xa][ib][t][id] : v;
                        out[xa][ib][xc][id] = v;
                    } else {
                        out[xa][ib][xc][id] = Kt[xa][ib][xc][id];
                    }
                }
}

// Primitive-pair data of a ket as the quartet loop reads it: exponent sum q,
// 1/q, centre Q, Q - C, and the prefactor K_cd (double, or float in the
// single-precision ket tables of the FP32 kernels).
template <typename T>
struct KetPrimT {
    T p, ip, Px, Py, Pz, QCx, QCy, QCz, K;
};
using KetPrim = KetPrimT<double>;

// Ket primitives stored as PrimPair records (the ket PrimPair keeps Q - C in
// its PA fields).
struct KetPrimsAoS {
    const PrimPair* pp;
    int np;
    HFX_HD HFX_INLINE int count() const { return np; }
    HFX_HD HFX_INLINE KetPrim get(int k) const {
        const PrimPair& q = pp[k];
        return KetPrim{q.p, q.ip, q.Px, q.Py, q.Pz, q.PAx, q.PAy, q.PAz, q.K};
    }
};

// Ket primitives of ket j of a sublist stored in the chunked planes of
// KetChunk (tab: the sublist's table, pitch: its plane length).
template <typename T>
struct KetPrimsSoA {
    const T* tab;
    int j;
    int np;
    int pitch;
    using C = typename KetChunkOf<T>::type;
    static constexpr int PER = KetChunk<T>::PER, NCH = KetChunk<T>::NCH;
    HFX_HD HFX_INLINE int count() const { return np; }
    HFX_HD HFX_INLINE KetPrimT<T> get(int k) const {
        const T* pk = tab + (size_t)k * KETPRIM_FIELDS * pitch;
        T f[8];
#pragma unroll
        for (int h = 0; h < NCH; h++) {
            const C c = reinterpret_cast<const C*>(pk + (size_t)h * pitch * PER)[j];
#pragma unroll
            for (int e = 0; e < PER; e++) f[h * PER + e] = c.v[e];
        }
        // slots: p, 1/p, Px, Py, Pz, K, QCx, QCy | QCz (plane 8)
        return KetPrimT<T>{f[0], f[1], f[2], f[3], f[4], f[6], f[7], pk[(size_t)8 * pitch + j], f[5]};
    }
    // exponent of the C primitive of primitive pair k
    HFX_HD HFX_INLINE T cexp(int k) const { return tab[((size_t)k * KETPRIM_FIELDS + 9) * pitch + j]; }
};

// Ket primitives as PrimPair records plus their exponents (derivative integrals).
struct KetPrimsAoSG : KetPrimsAoS {
    const PrimExp* pe;
    HFX_HD HFX_INLINE double cexp(int k) const { return pe[k].aa; }
};

// Rys rule of the FP32 kernels: 0 = the FP64 rules for every root count,
// 1 = the single-precision rules of rys.h for every root count, 2 = one and
// two roots from the Boys moments in FP64, three to five roots from the
// single-precision fits (relative error < 1.1e-8, below the rounding of the
// nodes and weights to float).
#ifndef HFX_FP32_RULES
#define HFX_FP32_RULES 2
#endif
// Taylor terms of the FP64 Boys moments in the FP32 kernels: 5 terms truncate
// at (1/64)^5/5! < 8e-12 relative, far below the float rounding of the nodes
// and weights they give (the FP64 kernels take all BOYS_K = 7).
#ifndef HFX_BOYS_K32
#define HFX_BOYS_K32 5
#endif
// (ss|ss) quartets take the compact one-root table (one 16-byte load and the
// downward recursion, rys_roots_one); other one-root classes the full Taylor
// table. Measured per class at 64 waters: (00|00) 6 % (FP64) and 9 % (FP32)
// faster with the compact table, (10|00) unchanged (its launch issues more
// arithmetic per table load, so the recursion's extra operations cancel the
// saved loads).
template <int LA, int LB, int LC, int LD>
HFX_HD constexpr bool boys_compact() { return LA + LB + LC + LD == 0; }

// Contracted integrals of the shell quartet (bra | ket), slice (sa, sc):
// eri[((a*NB + b)*NCt + c)*ND + d] for the thread's a- and c-components.
// bra: shell pair with l_A = LA, l_B = LB (primitive pairs prims[bra.p0 ..]);
// ket: l_C = LC, l_D = LD, C - D = (cdx, cdy, cdz), primitives from kp.
// Ket primitives form the outer loop (one per-lane load each); the bra's are
// the same for the whole wavefront, read as PrimPair or as its single-precision
// copy PrimPairF (BP).
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R, class KP, class BP>
HFX_HD HFX_INLINE void eri_quartet_k(const ShellPair& bra, const BP* __restrict__ prims, double cdx_,
                                     double cdy_, double cdz_, const KP& kp, int sa, int sc, const RysTables& tab,
                                     R (&eri)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT]) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    using PT = R;
    constexpr int NR = Q::NR;
#pragma unroll
    for (int i = 0; i < Q::NOUT; i++) eri[i] = R(0);
    // exponents of the selected a / c component (run-time, wavefront-uniform)
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
    const R abx = R(bra.ABx), aby = R(bra.ABy), abz = R(bra.ABz);
    const R cdx = R(cdx_), cdy = R(cdy_), cdz = R(cdz_);
    const int nkp = kp.count();
    for (int k1 = 0; k1 < nkp; k1++) {
        const auto Qp = kp.get(k1);
        const PT q = PT(Qp.p), iq = PT(Qp.ip);
        const PT kq = PT(Qp.K) * iq;   // the ket's factor K / q of the prefactor
        for (int i1 = 0; i1 < bra.np; i1++) {
            const BP P = hfx_ld_prim(prims + bra.p0 + i1);
            // primitive-level arithmetic in the kernel's type (the Rys rule of the
            // FP32 kernels in FP64 unless HFX_FP32_RULES)
            const PT p = PT(P.p), pq = p + q;
            const PT ip = PT(P.ip);
            const PT isq = hfx_rsqrt(pq);
            const PT ipq = isq * isq;
            const PT rho = p * q * ipq;
            const PT PQx = PT(P.Px - Qp.Px), PQy = PT(P.Py - Qp.Py), PQz = PT(P.Pz - Qp.Pz);
            const PT T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
            const PT pref = PT(P.Kp) * kq * isq;
            PT u[NR], w[NR];
            if constexpr (sizeof(PT) == 4 && (HFX_FP32_RULES == 1 || (HFX_FP32_RULES == 2 && NR >= 3))) {
                if constexpr (NR <= 2)
                    rys_roots_small_f<NR>(T, u, w, tab.boysf);
                else
                    rys_roots_f<NR>(T, u, w, tab);
            } else if constexpr (sizeof(PT) == 4) {
                // FP32 quartets: the Boys function and the Rys rule in FP64
                double ud[NR], wd[NR];
                if constexpr (boys_compact<LA, LB, LC, LD>())
                    rys_roots_one<HFX_BOYS_K32>(double(T), ud, wd, tab.boys1);
                else if constexpr (NR <= 2)
                    rys_roots_small<NR, HFX_BOYS_K32>(double(T), ud, wd, tab.boys);
                else
                    rys_roots<NR>(double(T), ud, wd, tab);
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    u[r] = PT(ud[r]);
                    w[r] = PT(wd[r]);
                }
            } else {
                if constexpr (boys_compact<LA, LB, LC, LD>())
                    rys_roots_one<BOYS_K>(T, u, w, tab.boys1);
                else if constexpr (NR <= 2)
                    rys_roots_small<NR>(T, u, w, tab.boys);
                else
                    rys_roots<NR>(T, u, w, tab);
            }
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const PT ur = u[r];
                const PT t1 = ur * rho;
                const R b00 = R(PT(0.5) * ur * ipq);
                const R b10 = R(PT(0.5) * ip * (PT(1) - t1 * ip));
                const R b01 = R(PT(0.5) * iq * (PT(1) - t1 * iq));
                const PT fp = t1 * ip, fq = t1 * iq;
                R Ix[Q::XA][LB + 1][Q::XC][LD + 1];
                R Iy[Q::XA][LB + 1][Q::XC][LD + 1];
                R Iz[Q::XA][LB + 1][Q::XC][LD + 1];
                rys_dir<LA, LB, LC, LD, SPA, SPC, R>(R(PT(P.PAx) - fp * PQx), R(PT(Qp.QCx) + fq * PQx), b10, b01, b00,
                                                     abx, cdx, R(1), sel_a[0], sel_c[0], Ix);
                rys_dir<LA, LB, LC, LD, SPA, SPC, R>(R(PT(P.PAy) - fp * PQy), R(PT(Qp.QCy) + fq * PQy), b10, b01, b00,
                                                     aby, cdy, R(1), sel_a[1], sel_c[1], Iy);
                rys_dir<LA, LB, LC, LD, SPA, SPC, R>(R(PT(P.PAz) - fp * PQz), R(PT(Qp.QCz) + fq * PQz), b10, b01, b00,
                                                     abz, cdz, R(w[r] * pref), sel_a[2], sel_c[2], Iz);
#pragma unroll
                for (int a = 0; a < Q::NAt; a++) {
                    const int ax = SPA ? 0 : cart_exp(LA, a, 0), ay = SPA ? 0 : cart_exp(LA, a, 1),
                              az = SPA ? 0 : cart_exp(LA, a, 2);
#pragma unroll
                    for (int b = 0; b < Q::NB; b++) {
                        const int bx = cart_exp(LB, b, 0), by = cart_exp(LB, b, 1), bz = cart_exp(LB, b, 2);
#pragma unroll
                        for (int c = 0; c < Q::NCt; c++) {
                            const int cx = SPC ? 0 : cart_exp(LC, c, 0), cy = SPC ? 0 : cart_exp(LC, c, 1),
                                      cz = SPC ? 0 : cart_exp(LC, c, 2);
#pragma unroll
                            for (int d = 0; d < Q::ND; d++) {
                                const int dx = cart_exp(LD, d, 0), dy = cart_exp(LD, d, 1), dz = cart_exp(LD, d, 2);
                                const R xy = Ix[ax][bx][cx][dx] * Iy[ay][by][cy][dy];
                                R& e = eri[((a * Q::NB + b) * Q::NCt + c) * Q::ND + d];
                                e = xy * Iz[az][bz][cz][dz] + e;
                            }
                        }
                    }
                }
            }
        }
    }
}

// The same for a ket given as a shell pair with PrimPair records.
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void eri_quartet(const ShellPair& bra, const ShellPair& ket, const PrimPair* __restrict__ prims,
                                   int sa, int sc, const RysTables& tab,
                                   R (&eri)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT]) {
    eri_quartet_k<LA, LB, LC, LD, SPA, SPC, R>(bra, prims, ket.ABx, ket.ABy, ket.ABz,
                                               KetPrimsAoS{prims + ket.p0, ket.np}, sa, sc, tab, eri);
}

// J/K digestion of one slice of a quartet: the contributions of one quartet
// to the six density-contracted blocks, accumulated in registers. The caller
// scales them by the quartet degeneracy factor s (1/2 for A==B, for C==D and
// for bra pair == ket pair) and flushes them into unsymmetrised buffers Jb, Kb
// (row index = first function), symmetrised afterwards (J = Jb + Jb^T, K = Kb + Kb^T):
//   Jb_ab += 2 s sum_cd (ab|cd) D_cd,  Jb_cd += 2 s sum_ab (ab|cd) D_ab,
//   Kb_ac += s (ab|cd) D_bd,  Kb_ad += s (ab|cd) D_bc,
//   Kb_bc += s (ab|cd) D_ad,  Kb_bd += s (ab|cd) D_ac.
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
struct JKAcc {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    R Jab[Q::NAt][Q::NB], Jcd[Q::NCt][Q::ND];
    R Kac[Q::NAt][Q::NCt], Kad[Q::NAt][Q::ND], Kbc[Q::NB][Q::NCt], Kbd[Q::NB][Q::ND];
};

template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void jk_zero(JKAcc<LA, LB, LC, LD, SPA, SPC, R>& acc) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
#pragma unroll
    for (int a = 0; a < Q::NAt; a++) {
#pragma unroll
        for (int b = 0; b < Q::NB; b++) acc.Jab[a][b] = R(0);
#pragma unroll
        for (int c = 0; c < Q::NCt; c++) acc.Kac[a][c] = R(0);
#pragma unroll
        for (int d = 0; d < Q::ND; d++) acc.Kad[a][d] = R(0);
    }
#pragma unroll
    for (int b = 0; b < Q::NB; b++) {
#pragma unroll
        for (int c = 0; c < Q::NCt; c++) acc.Kbc[b][c] = R(0);
#pragma unroll
        for (int d = 0; d < Q::ND; d++) acc.Kbd[b][d] = R(0);
    }
#pragma unroll
    for (int c = 0; c < Q::NCt; c++)
#pragma unroll
        for (int d = 0; d < Q::ND; d++) acc.Jcd[c][d] = R(0);
}

// The six density blocks one quartet slice is contracted with.
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
struct JKDens {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    R ab[Q::NAt][Q::NB], cd[Q::NCt][Q::ND];
    R ac[Q::NAt][Q::NCt], ad[Q::NAt][Q::ND], bc[Q::NB][Q::NCt], bd[Q::NB][Q::ND];
};

template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void jk_contract(const R (&eri)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT],
                                   const JKDens<LA, LB, LC, LD, SPA, SPC, R>& D,
                                   JKAcc<LA, LB, LC, LD, SPA, SPC, R>& acc) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int NAt = Q::NAt, NB = Q::NB, NCt = Q::NCt, ND = Q::ND;
#pragma unroll
    for (int a = 0; a < NAt; a++)
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int c = 0; c < NCt; c++)
#pragma unroll
                for (int d = 0; d < ND; d++) {
                    const R v = eri[((a * NB + b) * NCt + c) * ND + d];
                    acc.Jab[a][b] = v * D.cd[c][d] + acc.Jab[a][b];
                    acc.Jcd[c][d] = v * D.ab[a][b] + acc.Jcd[c][d];
                    acc.Kac[a][c] = v * D.bd[b][d] + acc.Kac[a][c];
                    acc.Kad[a][d] = v * D.bc[b][c] + acc.Kad[a][d];
                    acc.Kbc[b][c] = v * D.ad[a][d] + acc.Kbc[b][c];
                    acc.Kbd[b][d] = v * D.ac[a][c] + acc.Kbd[b][d];
                }
}

// fa/fc: first function of the thread's a / c slice (fa + sa when slicing).
template <int LA, int LB, int LC, int LD, bool SPA, bool SPC, typename R>
HFX_HD HFX_INLINE void jk_accumulate(const R (&eri)[QClass<LA, LB, LC, LD, SPA, SPC>::NOUT], int fa, int fb, int fc,
                                     int fd, const double* __restrict__ D, int n,
                                     JKAcc<LA, LB, LC, LD, SPA, SPC, R>& acc) {
    using Q = QClass<LA, LB, LC, LD, SPA, SPC>;
    constexpr int NAt = Q::NAt, NB = Q::NB, NCt = Q::NCt, ND = Q::ND;
    JKDens<LA, LB, LC, LD, SPA, SPC, R> dn;
#pragma unroll
    for (int a = 0; a < NAt; a++) {
#pragma unroll
        for (int b = 0; b < NB; b++) dn.ab[a][b] = R(D[(size_t)(fa + a) * n + fb + b]);
#pragma unroll
        for (int d = 0; d < ND; d++) dn.ad[a][d] = R(D[(size_t)(fa + a) * n + fd + d]);
#pragma unroll
        for (int c = 0; c < NCt; c++) dn.ac[a][c] = R(D[(size_t)(fa + a) * n + fc + c]);
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
#pragma unroll
        for (int d = 0; d < ND; d++) dn.bd[b][d] = R(D[(size_t)(fb + b) * n + fd + d]);
#pragma unroll
        for (int c = 0; c < NCt; c++) dn.bc[b][c] = R(D[(size_t)(fb + b) * n + fc + c]);
    }
#pragma unroll
    for (int c = 0; c < NCt; c++)
#pragma unroll
        for (int d = 0; d < ND; d++) dn.cd[c][d] = R(D[(size_t)(fc + c) * n + fd + d]);
    jk_contract<LA, LB, LC, LD, SPA, SPC, R>(eri, dn, acc);
}
