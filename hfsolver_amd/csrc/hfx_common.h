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
    double Kp;            // 2 pi^{5/2} K / p: the bra's factor of the ERI prefactor
};

// Single-precision copy of a primitive pair (same fields), read by the FP32
// kernels so that their primitive loops take the bra values without conversion.
struct PrimPairF {
    float p, ip, Px, Py, Pz, PAx, PAy, PAz, K, Kp;
};

// Exponents of a primitive pair (used by the derivative integrals).
struct PrimExp {
    double aa, ab;
};

static_assert(sizeof(ShellPair) == 80 && sizeof(PrimPair) == 80 && sizeof(PrimPairF) == 40,
              "record layouts read field by field below");

// Loads of read-only data whose address is the same for the whole wavefront:
// on the device through the constant address space, so that the compiler can
// use scalar loads (they stay vector loads when the address diverges).
#if defined(__HIP_DEVICE_COMPILE__)
#define HFX_AS4 __attribute__((address_space(4)))
#endif

HFX_HD HFX_INLINE int hfx_ld_int(const int* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const HFX_AS4 int*)p;
#else
    return *p;
#endif
}

HFX_HD HFX_INLINE double hfx_ld_dbl(const double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const HFX_AS4 double*)p;
#else
    return *p;
#endif
}

HFX_HD HFX_INLINE PrimPair hfx_ld_prim(const PrimPair* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    const HFX_AS4 double* d = (const HFX_AS4 double*)p;
    PrimPair r;
    r.p = d[0];
    r.ip = d[1];
    r.Px = d[2];
    r.Py = d[3];
    r.Pz = d[4];
    r.PAx = d[5];
    r.PAy = d[6];
    r.PAz = d[7];
    r.K = d[8];
    r.Kp = d[9];
    return r;
#else
    return *p;
#endif
}

HFX_HD HFX_INLINE PrimPairF hfx_ld_prim(const PrimPairF* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    const HFX_AS4 float* d = (const HFX_AS4 float*)p;
    PrimPairF r;
    r.p = d[0];
    r.ip = d[1];
    r.Px = d[2];
    r.Py = d[3];
    r.Pz = d[4];
    r.PAx = d[5];
    r.PAy = d[6];
    r.PAz = d[7];
    r.K = d[8];
    r.Kp = d[9];
    return r;
#else
    return *p;
#endif
}

// Primitive-pair record a kernel of arithmetic type R reads for its bra.
template <typename R>
struct BraPrimOf {
    using type = PrimPair;
};
template <>
struct BraPrimOf<float> {
    using type = PrimPairF;
};

HFX_HD HFX_INLINE PrimExp hfx_ld_pexp(const PrimExp* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    const HFX_AS4 double* d = (const HFX_AS4 double*)p;
    return PrimExp{d[0], d[1]};
#else
    return *p;
#endif
}

HFX_HD HFX_INLINE ShellPair hfx_ld_pair(const ShellPair* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    const HFX_AS4 int* q = (const HFX_AS4 int*)p;
    const HFX_AS4 double* d = (const HFX_AS4 double*)(q + 12);
    ShellPair r;
    r.A = q[0];
    r.B = q[1];
    r.fa = q[2];
    r.fb = q[3];
    r.p0 = q[4];
    r.np = q[5];
    r.cls = q[6];
    r.cpa = q[7];
    r.cpb = q[8];
    r.pad[0] = r.pad[1] = r.pad[2] = 0;
    r.ABx = d[0];
    r.ABy = d[1];
    r.ABz = d[2];
    r.Q = d[3];
    return r;
#else
    return *p;
#endif
}

// Fields of a ket primitive pair in a ket table: p, 1/p, P, P - C, K, and the
// exponent of the C primitive (derivative integrals), in this order.
static constexpr int KETPRIM_FIELDS = 10;

// Per primitive k of a ket sublist the table holds 10 planes of `pitch` words
// (pitch = the sublist's ket count rounded up to a multiple of 4), the fields
// stored in the slot order p, 1/p, P, K, (P - C)_x, (P - C)_y | (P - C)_z,
// C exponent: slots 0-7 in 16-byte chunks of two doubles or four floats,
// chunk h of ket j at word (10 k pitch) + (h pitch + j) * PER, slots 8 and 9
// as plain planes at words (10 k + 8) pitch + j and (10 k + 9) pitch + j.
// Consecutive kets (consecutive lanes) read consecutive words; an energy
// kernel loads a primitive with four (FP64) or two (FP32) 16-byte loads and
// one plain load per lane, and a class whose ket is an s-s pair (no P - C)
// with three (FP64) or two (FP32) 16-byte loads.
static constexpr int KETPRIM_SLOT[KETPRIM_FIELDS] = {0, 1, 2, 3, 4, 6, 7, 8, 5, 9};   // slot of each field
template <typename T>
struct KetChunk {
    static constexpr int PER = 16 / (int)sizeof(T);   // fields per chunk
    static constexpr int NCH = 8 / PER;               // chunks of the first eight fields
};
HFX_HD HFX_INLINE int ket_pitch(int nks) { return (nks + 3) & ~3; }
struct alignas(16) KetChunkD {
    double v[2];
};
struct alignas(16) KetChunkF {
    float v[4];
};
template <typename T>
struct KetChunkOf;
template <>
struct KetChunkOf<double> {
    using type = KetChunkD;
};
template <>
struct KetChunkOf<float> {
    using type = KetChunkF;
};


static constexpr double HFX_PI = 3.14159265358979323846;
static constexpr double HFX_SQRT_PI = 1.77245385090551602730;
// 2 pi^(5/2)
static constexpr double HFX_TWO_PI_52 = 34.98683665524972;
