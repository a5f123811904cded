// Rys quadrature nodes and weights from the generated piecewise-polynomial
// tables (rys_tables.inc, written by gen_rys_tables.py).
//
// Three to five roots: for T in [0, RYS_TMAX) the interval is
// it = floor(T / RYS_H), x = 2 (T - it H)/H - 1, and each u_i(T), w_i(T) is a
// polynomial of degree RYS_DEG in x, stored in powers of x (converted from a
// Chebyshev interpolant at high precision) and evaluated by Horner's rule; the
// single-precision tables of the FP32 kernels keep a truncated Chebyshev
// series. One and two roots: closed forms from the Boys moments F_0..F_3 of a
// Taylor table (BOYS_TAB: a 1/32 grid, BOYS_K terms). For T >= RYS_TMAX the
// rule is half-range Gauss-Hermite: u_i = x_i^2 / T, w_i = W_i / sqrt(T).
// sum_i w_i = F_0(T), the Boys function of hfsolver's src/special_functions.f90:189-248.
#pragma once

#include "hfx_common.h"
#include "rys_params.h"

// Pointers to the table data (device arrays in kernels, host arrays on the CPU).
struct RysTables {
    const double* poly;  // RYS_POLY
    const double* hu;    // RYS_HERM_U
    const double* hw;    // RYS_HERM_W
    const double* boys;  // BOYS_TAB (global memory, read through the caches)
    const float* boysf;  // BOYSF_TAB, single-precision Boys table of the FP32 kernels
    const float* chebf;  // RYS_CHEBF, single-precision truncated Chebyshev series of the FP32 kernels
    const double* boys1; // BOYS1_TAB, compact one-root table (F_top and e^{-T0} per grid point)
};

// 1/sqrt(x) for positive normal x (p + q, or T beyond the tables): on the device
// the hardware estimate refined by one Householder step of order 3,
// r <- r + r e (1/2 + 3/8 e) with e = 1 - x r^2, which takes the estimate's
// relative error eps to about (5/16) eps^3.
HFX_HD HFX_INLINE double hfx_rsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * r, r, 1.0);
    return fma(r * e, fma(e, 0.375, 0.5), r);
#else
    return 1.0 / std::sqrt(x);
#endif
}

// Single precision: the hardware reciprocal square root (1 ulp) on the device;
// every caller passes a positive normal number (p + q, or T beyond the table).
HFX_HD HFX_INLINE float hfx_rsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}

// i * j for i, j < 2^24 (table offsets): the 24-bit multiply on the device.
HFX_HD HFX_INLINE unsigned hfx_umul24(unsigned i, unsigned j) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(i, j);
#else
    return i * j;
#endif
}

// Square root in single precision: the hardware instruction (1 ulp) on the
// device; the argument (a discriminant of two distinct real roots) is positive.
HFX_HD HFX_INLINE float hfx_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}

// 1/x in single precision: the hardware reciprocal (1 ulp) on the device.
HFX_HD HFX_INLINE float hfx_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// 1/x for positive normal x: hardware reciprocal and two Newton steps on the
// device (within an ulp or two of the rounded quotient), a division on the host.
HFX_HD HFX_INLINE double hfx_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// Two-root rule for T >= BOYS_TMAX (moments (2m-1)!! F_0 / (2T)^m): nodes
// v / (2T) with v = 3 -+ sqrt(6), weights F_0 (sqrt(6) +- 2) / (2 sqrt(6)).
#define HFX_RYS2_V1 0.55051025721682190180   // 3 - sqrt(6)
#define HFX_RYS2_V2 5.44948974278317809820   // 3 + sqrt(6)
#define HFX_RYS2_W1 0.90824829046386301637   // (sqrt(6) + 2) / (2 sqrt(6))
#define HFX_RYS2_W2 0.09175170953613698363   // (sqrt(6) - 2) / (2 sqrt(6))

// One- and two-root rules in closed form from the Boys moments F_0..F_{2n-1}.
// F_m(T) for T < BOYS_TMAX: Taylor series about the nearest grid point T0,
//   F_m(T) = sum_{k < BOYS_K} F_{m+k}(T0) (T0 - T)^k / k!;
// beyond: F_0 = (1/2) sqrt(pi/T) (erfc(sqrt T) < 1e-16 F_0 there) and the
// upward recursion F_{m+1} = ((2m+1) F_m - e^{-T}) / (2T), in which e^{-T}
// (below 1.1e-20 for T >= 46) is under 1e-16 of (2m+1) F_m for m <= 2 and is
// left out, so that no exponential is evaluated; the rules then have the
// closed forms above (HFX_RYS2_*).
// n = 1: u = F_1/F_0, w = F_0. n = 2: the orthogonal quadratic of the moment
// sequence, D u^2 + A u + B with D = F_0 F_2 - F_1^2, A = F_1 F_2 - F_0 F_3,
// B = F_1 F_3 - F_2^2; with E = (-A + sqrt(A^2 - 4 D B))/2 the roots are
// u_2 = E / D and u_1 = B / E, the weights w_1 = (F_1 - u_2 F_0)/(u_1 - u_2)
// = E (F_1 D - E F_0) / (B D - E^2) and w_2 = F_0 - w_1. The three quotients
// share one reciprocal, 1 / (D E (B D - E^2)).
template <int NR, int K = BOYS_K>
HFX_HD HFX_INLINE void rys_roots_small(double T, double* u, double* w, const double* boys) {
    static_assert(NR == 1 || NR == 2, "closed form for one or two roots only");
    static_assert(K >= 2 && K <= BOYS_K, "Taylor terms beyond the table");
    constexpr int NM = 2 * NR;  // moments F_0 .. F_{2NR-1}
    double F[NM];
    if (T < BOYS_TMAX) {
        const unsigned i = (unsigned)(T * BOYS_INV_DT + 0.5);
        const double x = i * BOYS_DT - T;
        double xk[K];
#pragma unroll
        for (int k = 1; k < K; k++) xk[k] = x * (1.0 / k);
        const double* row = boys + hfx_umul24(i, BOYS_NCOL);
#pragma unroll
        for (int m = 0; m < NM; m++) {
            double acc = row[m + K - 1];
#pragma unroll
            for (int k = K - 2; k >= 0; k--) acc = row[m + k] + acc * xk[k + 1];
            F[m] = acc;
        }
        if constexpr (NR == 1) {
            u[0] = F[1] * hfx_rcp(F[0]);
            w[0] = F[0];
        }
    } else {
        const double rt = hfx_rsqrt(T);
        const double it2 = 0.5 * rt * rt;
        const double f0 = 0.5 * HFX_SQRT_PI * rt;
        if constexpr (NR == 1) {
            u[0] = it2;
            w[0] = f0;
        } else {
            // moments F_m = (2m-1)!! F_0 / (2T)^m: the orthogonal quadratic is
            // v^2 - 6v + 3 in v = 2T u, so u = (3 -+ sqrt 6)/(2T) and
            // w = F_0 (sqrt 6 +- 2)/(2 sqrt 6)
            u[0] = HFX_RYS2_V1 * it2;
            u[1] = HFX_RYS2_V2 * it2;
            w[0] = HFX_RYS2_W1 * f0;
            w[1] = HFX_RYS2_W2 * f0;
        }
        return;
    }
    if constexpr (NR == 2) {
        const double m0 = F[0], m1 = F[1], m2 = F[2], m3 = F[3];
        const double D = m0 * m2 - m1 * m1;
        const double A = m1 * m2 - m0 * m3;
        const double B = m1 * m3 - m2 * m2;
        const double E = 0.5 * (-A + sqrt(A * A - 4.0 * D * B));
        const double G = B * D - E * E;
        const double r = hfx_rcp(D * E * G);
        u[0] = B * D * G * r;
        u[1] = E * E * G * r;
        const double w1 = (m1 * D - E * m0) * E * D * E * r;
        w[0] = w1;
        w[1] = m0 - w1;
    }
}

// One-root rule from the compact table BOYS1_TAB: per grid point T0 the
// moment F_K(T0) and e^{-T0} in one 16-byte load (K = BOYS_K for the FP64
// kernels, BOYS1_KF for the FP32 kernels, which read the second half of the
// row); F_{K-1}(T0) .. F_0(T0) by the downward recursion
// F_m = (2 T0 F_{m+1} + e^{-T0}) / (2m + 1), stable in this direction; then
// F_0(T) and F_1(T) by the K-term Taylor series of rys_roots_small, and
// u = F_1 / F_0, w = F_0.
template <int K>
HFX_HD HFX_INLINE void rys_roots_one(double T, double* u, double* w, const double* boys1) {
    static_assert(K == BOYS_K || K == BOYS1_KF, "the table holds F_K for these K only");
    if (T < BOYS_TMAX) {
        const unsigned i = (unsigned)(T * BOYS_INV_DT + 0.5);
        const double T0 = i * BOYS_DT;
        const double x = T0 - T;
        const double* row = boys1 + 4 * i + (K == BOYS_K ? 0 : 2);
        const double e = row[1];
        double Fm[K + 1];
        Fm[K] = row[0];
#pragma unroll
        for (int m = K - 1; m >= 0; m--) Fm[m] = fma(2.0 * T0, Fm[m + 1], e) * (1.0 / (2 * m + 1));
        double xk[K];
#pragma unroll
        for (int k = 1; k < K; k++) xk[k] = x * (1.0 / k);
        double f0 = Fm[K - 1], f1 = Fm[K];
#pragma unroll
        for (int k = K - 2; k >= 0; k--) {
            f0 = Fm[k] + f0 * xk[k + 1];
            f1 = Fm[k + 1] + f1 * xk[k + 1];
        }
        u[0] = f1 * hfx_rcp(f0);
        w[0] = f0;
    } else {
        const double rt = hfx_rsqrt(T);
        u[0] = 0.5 * rt * rt;
        w[0] = 0.5 * HFX_SQRT_PI * rt;
    }
}

// The same rules in single precision for the FP32 kernels, from BOYSF_TAB
// (BOYSF_K Taylor terms, relative error below float rounding); beyond
// BOYS_TMAX the same asymptotic closed forms.
template <int NR>
HFX_HD HFX_INLINE void rys_roots_small_f(float T, float* u, float* w, const float* boysf) {
    static_assert(NR == 1 || NR == 2, "closed form for one or two roots only");
    constexpr int NM = 2 * NR;
    float F[NM];
    if (T < (float)BOYS_TMAX) {
        const int i = (int)(T * (float)BOYS_INV_DT + 0.5f);
        const float x = (float)i * (float)BOYS_DT - T;
        float xk[BOYSF_K];
#pragma unroll
        for (int k = 1; k < BOYSF_K; k++) xk[k] = x * (1.0f / k);
        const float* row = boysf + i * BOYSF_NCOL;
#pragma unroll
        for (int m = 0; m < NM; m++) {
            float acc = row[m + BOYSF_K - 1];
#pragma unroll
            for (int k = BOYSF_K - 2; k >= 0; k--) acc = row[m + k] + acc * xk[k + 1];
            F[m] = acc;
        }
        if constexpr (NR == 1) {
            u[0] = F[1] * hfx_rcp(F[0]);
            w[0] = F[0];
        }
    } else {
        // the asymptotic rules of rys_roots_small (no moments: their products
        // would leave the single-precision range for large T)
        const float rt = hfx_rsqrt(T);
        const float it2 = 0.5f * rt * rt;
        const float f0 = 0.5f * (float)HFX_SQRT_PI * rt;
        if constexpr (NR == 1) {
            u[0] = it2;
            w[0] = f0;
        } else {
            u[0] = (float)HFX_RYS2_V1 * it2;
            u[1] = (float)HFX_RYS2_V2 * it2;
            w[0] = (float)HFX_RYS2_W1 * f0;
            w[1] = (float)HFX_RYS2_W2 * f0;
        }
        return;
    }
    if constexpr (NR == 2) {
        const float m0 = F[0], m1 = F[1], m2 = F[2], m3 = F[3];
        const float id = hfx_rcp(m0 * m2 - m1 * m1);
        const float a = (m1 * m2 - m0 * m3) * id;
        const float b = (m1 * m3 - m2 * m2) * id;
        const float u2 = 0.5f * (-a + hfx_sqrt(a * a - 4.0f * b));
        const float u1 = b * hfx_rcp(u2);
        const float w1 = (m1 - u2 * m0) * hfx_rcp(u1 - u2);
        u[0] = u1;
        w[0] = w1;
        u[1] = u2;
        w[1] = m0 - w1;
    }
}

// Offset (in doubles) of the n-root block of RYS_POLY.
HFX_HD constexpr int rys_offset(int n) {
    int o = 0;
    for (int m = 1; m < n; m++) o += RYS_NINT * (RYS_DEG + 1) * 2 * m;
    return o;
}

template <int NR>
HFX_HD HFX_INLINE void rys_roots(double T, double* u, double* w, const RysTables& tab) {
    static_assert(NR >= 1 && NR <= RYS_NMAX, "root count outside the tables");
    if (T >= RYS_TMAX) {
        const double it = 1.0 / T;
        const double st = sqrt(it);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            u[i] = tab.hu[(NR - 1) * RYS_NMAX + i] * it;
            w[i] = tab.hw[(NR - 1) * RYS_NMAX + i] * st;
        }
        return;
    }
    int itv = (int)(T * (1.0 / RYS_H));
    if (itv > RYS_NINT - 1) itv = RYS_NINT - 1;
    const double x = 2.0 * (T - itv * RYS_H) * (1.0 / RYS_H) - 1.0;
    const double* c = tab.poly + rys_offset(NR) + itv * (RYS_DEG + 1) * 2 * NR;
    double acc[2 * NR];
#pragma unroll
    for (int f = 0; f < 2 * NR; f++) acc[f] = c[RYS_DEG * 2 * NR + f];
#pragma unroll
    for (int k = RYS_DEG - 1; k >= 0; k--) {
#pragma unroll
        for (int f = 0; f < 2 * NR; f++) acc[f] = fma(acc[f], x, c[k * 2 * NR + f]);
    }
#pragma unroll
    for (int i = 0; i < NR; i++) {
        u[i] = acc[2 * i];
        w[i] = acc[2 * i + 1];
    }
}

// Offset (in floats) of the n-root block of RYS_CHEBF.
HFX_HD constexpr int rys_offset_f(int n) {
    int o = 0;
    for (int m = 1; m < n; m++) o += RYS_NINT * (RYS_DEGF + 1) * 2 * m;
    return o;
}

// Single-precision nodes and weights for the FP32 kernels: the same piecewise
// Chebyshev series truncated after degree RYS_DEGF (relative truncation error
// below 1.1e-8), float Clenshaw; Gauss-Hermite beyond RYS_TMAX.
template <int NR>
HFX_HD HFX_INLINE void rys_roots_f(float T, float* u, float* w, const RysTables& tab) {
    static_assert(NR >= 1 && NR <= RYS_NMAX, "root count outside the tables");
    if (T >= (float)RYS_TMAX) {
        const float it = hfx_rcp(T);
        const float st = hfx_rsqrt(T);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            u[i] = (float)tab.hu[(NR - 1) * RYS_NMAX + i] * it;
            w[i] = (float)tab.hw[(NR - 1) * RYS_NMAX + i] * st;
        }
        return;
    }
    int itv = (int)(T * (float)(1.0 / RYS_H));
    if (itv > RYS_NINT - 1) itv = RYS_NINT - 1;
    const float x = 2.0f * (T - (float)itv * (float)RYS_H) * (float)(1.0 / RYS_H) - 1.0f;
    const float x2 = 2.0f * x;
    const float* c = tab.chebf + rys_offset_f(NR) + itv * (RYS_DEGF + 1) * 2 * NR;
    float b1[2 * NR], b2[2 * NR];
#pragma unroll
    for (int f = 0; f < 2 * NR; f++) {
        b1[f] = 0.0f;
        b2[f] = 0.0f;
    }
#pragma unroll
    for (int k = RYS_DEGF; k >= 1; k--) {
#pragma unroll
        for (int f = 0; f < 2 * NR; f++) {
            const float t = c[k * 2 * NR + f] + x2 * b1[f] - b2[f];
            b2[f] = b1[f];
            b1[f] = t;
        }
    }
#pragma unroll
    for (int i = 0; i < NR; i++) {
        u[i] = c[2 * i] + x * b1[2 * i] - b2[2 * i];
        w[i] = c[2 * i + 1] + x * b1[2 * i + 1] - b2[2 * i + 1];
    }
}
