// Rys quadrature nodes and weights from the generated piecewise-Chebyshev
// tables (rys_tables.inc, written by gen_rys_tables.py).
//
// For T in [0, RYS_TMAX): interval it = floor(T / RYS_H), x = 2 (T - it H)/H - 1,
// and each u_i(T), w_i(T) is a Chebyshev series of degree RYS_DEG in x,
// evaluated by Clenshaw's recurrence. For T >= RYS_TMAX the rule is
// half-range Gauss-Hermite: u_i = x_i^2 / T, w_i = W_i / sqrt(T).
// sum_i w_i = F_0(T), the Boys function of hfsolver's src/special_functions.f90:189-248.
#pragma once

#include "hfx_common.h"
#include "rys_params.h"

// Pointers to the table data (device arrays in kernels, host arrays on the CPU).
struct RysTables {
    const double* cheb;  // RYS_CHEB
    const double* hu;    // RYS_HERM_U
    const double* hw;    // RYS_HERM_W
};

// Offset (in doubles) of the n-root block of RYS_CHEB.
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
    const double x2 = 2.0 * x;
    const double* c = tab.cheb + rys_offset(NR) + itv * (RYS_DEG + 1) * 2 * NR;
    double b1[2 * NR], b2[2 * NR];
#pragma unroll
    for (int f = 0; f < 2 * NR; f++) {
        b1[f] = 0.0;
        b2[f] = 0.0;
    }
#pragma unroll
    for (int k = RYS_DEG; k >= 1; k--) {
#pragma unroll
        for (int f = 0; f < 2 * NR; f++) {
            const double t = c[k * 2 * NR + f] + x2 * b1[f] - b2[f];
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
