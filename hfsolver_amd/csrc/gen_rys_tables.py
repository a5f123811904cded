"""Generate the Rys quadrature tables used by the HIP kernels and host code.

Rys quadrature replaces the Boys-function sum of an electron-repulsion (or
nuclear-attraction) integral by an n-point Gauss rule for the weight
exp(-T t^2) on t in [0, 1], in the variable u = t^2:

    int_0^1 exp(-T t^2) g(t^2) dt = sum_i w_i(T) g(u_i(T))    (deg g <= 2n-1)

The moments of that weight are the Boys functions F_k(T) = int_0^1 t^2k
exp(-T t^2) dt (the function the reference evaluates in
src/special_functions.f90:189-248), so sum_i w_i = F_0(T).

Method (all published, nothing tabulated by hand):
  * moments F_0..F_{2n-1}(T) from the confluent hypergeometric form
    F_k(T) = 1F1(k+1/2; k+3/2; -T) / (2k+1), in 60-digit arithmetic (mpmath);
  * recurrence coefficients of the monic orthogonal polynomials by Gautschi's
    (ordinary) Chebyshev algorithm - ill-conditioned in double precision, but
    harmless at 60 digits;
  * nodes and weights from the Jacobi matrix (Golub-Welsch);
  * each u_i(T), w_i(T) is fitted on [0, TMAX) by piecewise Chebyshev
    interpolation (NINT intervals of width H, degree DEG), and the series is
    rewritten in powers of the interval variable x in [-1, 1] (in 60-digit
    arithmetic, then rounded): the kernels evaluate it by Horner's rule, one
    multiply-add per degree; the Chebyshev coefficients decay fast enough that
    sum_k |a_k| stays below 1.2 |f(x)| on every interval, so the power form
    loses at most a few ulps;
  * for T >= TMAX the weight exp(-T t^2) is negligible at t = 1 and the rule
    becomes half-range Gauss-Hermite: u_i = x_i^2 / T, w_i = W_i / sqrt(T)
    with x_i > 0 the positive nodes of the 2n-point Gauss-Hermite rule and W_i
    its weights.

Output: rys_params.h (sizes) and rys_tables.inc (C/HIP source, doubles), layout

    RYS_POLY[n][interval][power k][2*i + {0: u, 1: w}]   (flattened, see offsets)

so that one lane evaluating root i of one interval reads 2*DEG+2 doubles
that sit in one contiguous stretch per coefficient row.

Run:  python hfsolver_amd/csrc/gen_rys_tables.py   (about a minute, 8 cores)
"""
import os
import sys
from multiprocessing import Pool

import mpmath as mp

NMAX = 5          # enough for (dd|dd) energies and their first derivatives (L <= 9)
H = 0.125         # interval width in T
NINT = 480        # intervals on [0, TMAX); exp(-T) T^(m-1/2)/Gamma(m+1/2) < 1e-16 at T = 60 for m <= 9
TMAX = H * NINT
DEG = 7           # degree per interval: the Chebyshev tail beyond it is < 2e-17 of |f| on every interval
DEGF = 6          # single-precision copy: the series truncated after degree 6
                  # (dropped coefficients sum to < 1.1e-8 of |c_0| on every interval)
DPS = 60
# Boys-function Taylor table for the one- and two-root rules (closed form from
# F_0..F_3): F_m(T) = sum_{k<BOYS_K} F_{m+k}(T0) (T0-T)^k / k!, T0 on a grid of
# step BOYS_DT = 1/32, |T - T0| <= 1/64, truncation < (1/64)^7/7! ~ 5e-17
# relative (grid 1/8 with 9 terms has the same accuracy; the denser grid needs
# fewer loads and multiply-adds per evaluation). Per grid point the table
# holds F_0 .. F_{BOYS_MMAX + BOYS_K - 1} = F_0 .. F_9.
BOYS_DT = 0.03125
BOYS_TMAX = 46.0
BOYS_K = 7
BOYS_MMAX = 3                      # highest order needed (two roots: F_0..F_3)
BOYS_NCOL = BOYS_MMAX + BOYS_K     # F_0 .. F_9 per grid point
BOYS_NPT = int(BOYS_TMAX / BOYS_DT) + 1
# Compact one-root table on the same grid: per grid point F_top(T0) and e^{-T0}
# for top = BOYS_K (the FP64 kernels' Taylor terms) and top = BOYS1_KF (the
# FP32 kernels'), so that a one-root rule reads 16 bytes and recovers
# F_{top-1}(T0) .. F_0(T0) by the (stable) downward recursion
# F_m = (2 T0 F_{m+1} + e^{-T0}) / (2m + 1).
BOYS1_KF = 5
# single-precision copy on the same grid for the FP32 kernels (used only by a
# build with HFX_FP32_RULES=1): BOYSF_K terms, truncation < (1/64)^5/5! ~ 8e-12
# relative, below float rounding
BOYSF_K = 5
BOYSF_NCOL = BOYS_MMAX + BOYSF_K   # F_0 .. F_7 per grid point

mp.mp.dps = DPS


def boys_moments(T, kmax):
    T = mp.mpf(T)
    return [mp.hyp1f1(k + mp.mpf(1) / 2, k + mp.mpf(3) / 2, -T) / (2 * k + 1)
            for k in range(kmax + 1)]


def rys_exact(n, T):
    """Nodes u_i and weights w_i (ascending u) of the n-point Rys rule at T."""
    mu = boys_moments(T, 2 * n - 1)
    # Gautschi's Chebyshev algorithm on ordinary moments.
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mu)
    alpha[0] = mu[1] / mu[0]
    beta[0] = mu[0]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1]
        beta[k] = sig_new[k] / sig[k - 1]
        sig_prev, sig = sig, sig_new
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(J)
    pairs = sorted((E[i], mu[0] * Q[0, i] ** 2) for i in range(n))
    return [p[0] for p in pairs], [p[1] for p in pairs]


def hermite_half(n):
    """Positive nodes x_i^2 and weights of the 2n-point Gauss-Hermite rule."""
    m = 2 * n
    J = mp.matrix(m, m)
    for i in range(m - 1):
        J[i, i + 1] = J[i + 1, i] = mp.sqrt(mp.mpf(i + 1) / 2)
    E, Q = mp.eigsy(J)
    out = sorted((E[i] ** 2, mp.sqrt(mp.pi) * Q[0, i] ** 2) for i in range(m) if E[i] > 0)
    return [o[0] for o in out], [o[1] for o in out]


def fit_interval(args):
    n, it = args
    mp.mp.dps = DPS
    a = mp.mpf(it) * H
    vals = []  # vals[j] = (u list, w list) at node j
    xs = []
    for j in range(DEG + 1):
        x = mp.cos(mp.pi * (j + mp.mpf(1) / 2) / (DEG + 1))
        xs.append(x)
        T = a + H * (x + 1) / 2
        vals.append(rys_exact(n, T))
    cheb = []  # cheb[k][2*i + s], 60 digits
    for k in range(DEG + 1):
        row = []
        for i in range(n):
            for s in range(2):
                acc = mp.mpf(0)
                for j in range(DEG + 1):
                    acc += vals[j][s][i] * mp.cos(mp.pi * k * (j + mp.mpf(1) / 2) / (DEG + 1))
                acc *= mp.mpf(2) / (DEG + 1)
                if k == 0:
                    acc /= 2
                row.append(acc)
        cheb.append(row)
    # power form: coefficients of the Chebyshev polynomials T_k(x) by
    # T_{k+1} = 2x T_k - T_{k-1}, then a_m = sum_k c_k [x^m] T_k
    tk = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
    for k in range(2, DEG + 1):
        nxt = [mp.mpf(0)] * (k + 1)
        for m, v in enumerate(tk[k - 1]):
            nxt[m + 1] += 2 * v
        for m, v in enumerate(tk[k - 2]):
            nxt[m] -= v
        tk.append(nxt)
    poly = [[float(sum(cheb[k][f] * (tk[k][m] if m < len(tk[k]) else 0) for k in range(DEG + 1)))
             for f in range(2 * n)] for m in range(DEG + 1)]
    return n, it, [[float(v) for v in row] for row in cheb], poly


def main():
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rys_tables.inc")
    jobs = [(n, it) for n in range(1, NMAX + 1) for it in range(NINT)]
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(fit_interval, jobs)
    table = {(n, it): c for n, it, c, p in res}
    ptable = {(n, it): p for n, it, c, p in res}
    o = 0
    for n in range(1, NMAX + 1):
        o += NINT * (DEG + 1) * 2 * n
    hdr = ["// Generated by gen_rys_tables.py - do not edit.",
           "// Piecewise-polynomial Rys nodes/weights on T in [0, %g), Gauss-Hermite beyond." % TMAX,
           "#pragma once",
           "#define RYS_NMAX %d" % NMAX,
           "#define RYS_NINT %d" % NINT,
           "#define RYS_DEG %d" % DEG,
           "#define RYS_H %.17g" % H,
           "#define RYS_TMAX %.17g" % TMAX,
           "#define RYS_POLY_SIZE %d" % o,
           "#define RYS_DEGF %d" % DEGF,
           "#define RYS_CHEBF_SIZE %d" % (o // (DEG + 1) * (DEGF + 1)),
           "#define BOYS_DT %.17g" % BOYS_DT,
           "#define BOYS_INV_DT %.17g" % (1 / BOYS_DT),
           "#define BOYS_TMAX %.17g" % BOYS_TMAX,
           "#define BOYS_K %d" % BOYS_K,
           "#define BOYS_NCOL %d" % BOYS_NCOL,
           "#define BOYS_NPT %d" % BOYS_NPT,
           "#define BOYS_SIZE %d" % (BOYS_NCOL * BOYS_NPT),
           "#define BOYS1_KF %d" % BOYS1_KF,
           "#define BOYS1_SIZE %d" % (4 * BOYS_NPT),
           "#define BOYSF_K %d" % BOYSF_K,
           "#define BOYSF_NCOL %d" % BOYSF_NCOL,
           "#define BOYSF_SIZE %d" % (BOYSF_NCOL * BOYS_NPT)]
    with open(os.path.join(os.path.dirname(out), "rys_params.h"), "w") as f:
        f.write("\n".join(hdr) + "\n")
    lines = ["// Generated by gen_rys_tables.py - do not edit. Data of rys_params.h;",
             "// define RYS_TABLE_QUALIFIER before including."]
    lines.append("RYS_TABLE_QUALIFIER double RYS_POLY[RYS_POLY_SIZE] = {")
    for n in range(1, NMAX + 1):
        lines.append("// n = %d" % n)
        for it in range(NINT):
            for row in ptable[(n, it)]:
                lines.append("  " + ", ".join("%.17e" % v for v in row) + ",")
    lines.append("};")
    # single-precision truncation for the FP32 kernels, same layout with RYS_DEGF + 1 coefficients
    lines.append("RYS_TABLE_QUALIFIER float RYS_CHEBF[RYS_CHEBF_SIZE] = {")
    for n in range(1, NMAX + 1):
        lines.append("// n = %d" % n)
        for it in range(NINT):
            for row in table[(n, it)][:DEGF + 1]:
                lines.append("  " + ", ".join("%.9ef" % float(v) for v in row) + ",")
    lines.append("};")
    # asymptotic Gauss-Hermite data: RYS_HERM_U[n-1][i], RYS_HERM_W[n-1][i]
    hu, hw = [], []
    for n in range(1, NMAX + 1):
        r, w = hermite_half(n)
        hu.append([float(x) for x in r] + [0.0] * (NMAX - n))
        hw.append([float(x) for x in w] + [0.0] * (NMAX - n))
    lines.append("RYS_TABLE_QUALIFIER double RYS_HERM_U[RYS_NMAX * RYS_NMAX] = {")
    for row in hu:
        lines.append("  " + ", ".join("%.17e" % v for v in row) + ",")
    lines.append("};")
    lines.append("RYS_TABLE_QUALIFIER double RYS_HERM_W[RYS_NMAX * RYS_NMAX] = {")
    for row in hw:
        lines.append("  " + ", ".join("%.17e" % v for v in row) + ",")
    lines.append("};")
    # Boys table: BOYS_TAB[i * BOYS_NCOL + m] = F_m(i * BOYS_DT)
    lines.append("RYS_TABLE_QUALIFIER double BOYS_TAB[BOYS_SIZE] = {")
    fms = [boys_moments(mp.mpf(i) * mp.mpf(BOYS_DT), BOYS_NCOL - 1) for i in range(BOYS_NPT)]
    for fm in fms:
        lines.append("  " + ", ".join("%.17e" % float(v) for v in fm) + ",")
    lines.append("};")
    # compact one-root table: BOYS1_TAB[4 i .. 4 i + 3] = F_{BOYS_K}(T0), e^{-T0}, F_{BOYS1_KF}(T0), e^{-T0}
    lines.append("RYS_TABLE_QUALIFIER double BOYS1_TAB[BOYS1_SIZE] = {")
    for i in range(BOYS_NPT):
        T0 = mp.mpf(i) * mp.mpf(BOYS_DT)
        top = boys_moments(T0, max(BOYS_K, BOYS1_KF))
        e = mp.exp(-T0)
        lines.append("  %.17e, %.17e, %.17e, %.17e," % (float(top[BOYS_K]), float(e), float(top[BOYS1_KF]), float(e)))
    lines.append("};")
    # its single-precision copy: BOYSF_TAB[i * BOYSF_NCOL + m] = F_m(i * BOYS_DT), m < BOYSF_NCOL
    lines.append("RYS_TABLE_QUALIFIER float BOYSF_TAB[BOYSF_SIZE] = {")
    for fm in fms:
        lines.append("  " + ", ".join("%.9ef" % float(v) for v in fm[:BOYSF_NCOL]) + ",")
    lines.append("};")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", out, "doubles:", o)


if __name__ == "__main__":
    sys.exit(main())
