// Shared definitions of the direct J/K engine (host and device).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HFX_HD __host__ __device__
#define HFX_INLINE __forceinline__
#else
#define HFX_HD
#define HFX_INLINE inline
#endif

#define HFX_LMAX 2          // s, p, d shells (6-31G**)
#define HFX_NPAIRCLASS 6    // (ss) (ps) (pp) (ds) (dp) (dd)

// Number of Cartesian functions of a shell of angular momentum L.
HFX_HD constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

// Exponents (lx, ly, lz) of component i of a Cartesian shell of angular
// momentum L, in hfsolver's order: for k = 0..L, n = 0..k: (L-k, k-n, n)
// (src/basis.f90:371-409).
HFX_HD constexpr int cart_exp(int L, int i, int dim) {
    int idx = 0;
    for (int k = 0; k <= L; k++)
        for (int n = 0; n <= k; n++) {
            if (idx == i) return dim == 0 ? L - k : (dim == 1 ? k - n : n);
            idx++;
        }
    return -1;
}

// Pair classes in a fixed order; a pair (A,B) is stored with l_A >= l_B.
HFX_HD constexpr int pair_class(int la, int lb) {
    return la == 0 ? 0 : (la == 1 ? (lb == 0 ? 1 : 2) : (lb == 0 ? 3 : (lb == 1 ? 4 : 5)));
}
HFX_HD constexpr int pair_class_la(int c) { return c == 0 ? 0 : (c <= 2 ? 1 : 2); }
HFX_HD constexpr int pair_class_lb(int c) { return c == 0 ? 0 : (c == 1 ? 0 : (c == 2 ? 1 : c - 3)); }

// One shell pair (A, B) with l_A >= l_B. 80 bytes.
struct ShellPair {
    int A, B;      // shell indices
    int fa, fb;    // first (scaled-basis) function of A and of B
    int p0, np;    // primitive pairs [p0, p0 + np)
    int cls;       // pair class
    int cpa, cpb;  // first function of A / B among the functions of shells of the same L
    int pad[3];
    double ABx, ABy, ABz;  // A - B
    double Q;              // Schwarz bound sqrt(max |(ab|ab)|)
};

// One primitive pair of a shell pair. 80 bytes.
struct PrimPair {
    double p;             // alpha_a + alpha_b
    double ip;            // 1 / p
    double Px, Py, Pz;    // (alpha_a A + alpha_b B) / p
    double PAx, PAy, PAz; // P - A
    double K;             // c_a c_b exp(-alpha_a alpha_b |A-B|^2 / p), shell-level coefficients
    double pad;
};

// Fields of a ket primitive pair in a field-major ket table: p, 1/p, P, P - C, K.
static constexpr int KETPRIM_FIELDS = 9;

// Exponents of a primitive pair (used by the derivative integrals).
struct PrimExp {
    double aa, ab;
};

static constexpr double HFX_PI = 3.14159265358979323846;
// 2 pi^(5/2)
static constexpr double HFX_TWO_PI_52 = 34.98683665524972;
