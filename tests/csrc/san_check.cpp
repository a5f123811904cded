// Sanitizer driver (CPU only, test infrastructure): runs the engine's kernel
// templates instantiated for the host (eri_check.cpp: Rys rules, quartet
// integrals, derivative integrals) over every class combination, both
// arithmetic types and the edges of the rule tables, in a program built with
// AddressSanitizer and UndefinedBehaviorSanitizer. Any out-of-bounds table
// read, overflow or other undefined behaviour aborts with a report; the
// program also checks that every result is finite and that FP32 blocks of
// magnitude above 1e-10 stay within 1e-3 of the FP64 ones. Exit status 0 and
// "san_check: ok" on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
int chk_eri_quartet(int la, int lb, int lc, int ld, const double* A, const double* B, const double* C,
                    const double* D, int na, const double* aa, const double* ca, int nb, const double* ab,
                    const double* cb, int nc, const double* ac, const double* cc, int nd, const double* ad,
                    const double* cd, int fp32, double* out);
int chk_grad_quartet(int la, int lb, int lc, int ld, const double* A, const double* B, const double* C,
                     const double* D, int na, const double* aa, const double* ca, int nb, const double* ab,
                     const double* cb, int nc, const double* ac, const double* cc, int nd, const double* ad,
                     const double* cd, const double* gam, int fp32, double* g);
double chk_rys(int nr, double T, double* u, double* w);
}

namespace {

int ncart(int l) { return (l + 1) * (l + 2) / 2; }

int fails = 0;

void check(bool ok, const char* what, int a, int b, int c, int d) {
    if (ok) return;
    if (fails < 20) std::printf("san_check: FAILED %s (%d%d|%d%d)\n", what, a, b, c, d);
    fails++;
}

struct Shell {
    int L;
    double cen[3];
    std::vector<double> a, c;
};

Shell random_shell(std::mt19937_64& rng, int L, int np) {
    std::uniform_real_distribution<double> pos(-2.0, 2.0), lexp(-1.5, 3.0), coef(0.1, 1.0);
    Shell s;
    s.L = L;
    for (double& x : s.cen) x = pos(rng);
    for (int i = 0; i < np; i++) {
        s.a.push_back(std::pow(10.0, lexp(rng)));
        s.c.push_back(coef(rng));
    }
    return s;
}

}  // namespace

int main() {
    // Rys rules over [0, 100] (grid and table edges) and far beyond
    std::vector<double> Ts = {0.0, 1e-300, 1e-16, 1e-8, 45.999999, 46.0, 46.000001, 59.999999, 60.0, 60.000001,
                              1e3, 1e6, 1e12};
    for (int k = 0; k <= 100 * 64; k++) Ts.push_back(k / 64.0);
    const int codes[] = {1, 2, 3, 4, 5, -1, -2, -11, -12, -21, -22, -23, -24, -25};
    for (double T : Ts)
        for (int code : codes) {
            double u[5] = {0}, w[5] = {0};
            const int nr = code > 0 ? code : (code > -10 ? -code : (code > -20 ? -code - 10 : -code - 20));
            chk_rys(code, T, u, w);
            bool ok = true;
            for (int i = 0; i < nr; i++) ok = ok && std::isfinite(u[i]) && std::isfinite(w[i]) && u[i] >= 0 && w[i] >= 0;
            check(ok, "finite Rys rule", code, 0, (int)T, 0);
        }
    // every class combination, contracted and uncontracted shells, FP64 and FP32
    std::mt19937_64 rng(2024);
    for (int rep = 0; rep < 4; rep++)
        for (int la = 0; la <= 2; la++)
            for (int lb = 0; lb <= la; lb++)
                for (int lc = 0; lc <= 2; lc++)
                    for (int ld = 0; ld <= lc; ld++) {
                        const int X = la * (la + 1) / 2 + lb, Y = lc * (lc + 1) / 2 + ld;
                        if (X < Y) continue;
                        const int np = rep % 2 ? 1 : 3;
                        Shell A = random_shell(rng, la, np), B = random_shell(rng, lb, np),
                              C = random_shell(rng, lc, rep == 3 ? 6 : np), D = random_shell(rng, ld, np);
                        if (rep == 2) C.cen[0] = A.cen[0], C.cen[1] = A.cen[1], C.cen[2] = A.cen[2];  // T = 0 limit
                        const int n = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
                        std::vector<double> o64(n), o32(n), gam(n, 0.0);
                        for (int i = 0; i < n; i++) gam[i] = std::sin(1.0 + i);
                        int rc = chk_eri_quartet(la, lb, lc, ld, A.cen, B.cen, C.cen, D.cen, (int)A.a.size(), A.a.data(),
                                                 A.c.data(), (int)B.a.size(), B.a.data(), B.c.data(), (int)C.a.size(),
                                                 C.a.data(), C.c.data(), (int)D.a.size(), D.a.data(), D.c.data(), 0,
                                                 o64.data());
                        rc |= chk_eri_quartet(la, lb, lc, ld, A.cen, B.cen, C.cen, D.cen, (int)A.a.size(), A.a.data(),
                                              A.c.data(), (int)B.a.size(), B.a.data(), B.c.data(), (int)C.a.size(),
                                              C.a.data(), C.c.data(), (int)D.a.size(), D.a.data(), D.c.data(), 1,
                                              o32.data());
                        check(rc == 0, "quartet accepted", la, lb, lc, ld);
                        double m = 0, d = 0;
                        bool fin = true;
                        for (int i = 0; i < n; i++) {
                            fin = fin && std::isfinite(o64[i]) && std::isfinite(o32[i]);
                            m = std::fmax(m, std::fabs(o64[i]));
                            d = std::fmax(d, std::fabs(o64[i] - o32[i]));
                        }
                        check(fin, "finite integrals", la, lb, lc, ld);
                        // blocks of the magnitudes the engine evaluates in FP32 (bounds 1e-12 .. 1e-5; far
                        // smaller values of these random shells underflow single precision)
                        check(m < 1e-10 || d <= 1e-3 * m, "FP32 close to FP64", la, lb, lc, ld);
                        for (int fp32 = 0; fp32 < 2; fp32++) {
                            double g[9] = {0};
                            rc = chk_grad_quartet(la, lb, lc, ld, A.cen, B.cen, C.cen, D.cen, (int)A.a.size(),
                                                  A.a.data(), A.c.data(), (int)B.a.size(), B.a.data(), B.c.data(),
                                                  (int)C.a.size(), C.a.data(), C.c.data(), (int)D.a.size(),
                                                  D.a.data(), D.c.data(), gam.data(), fp32, g);
                            bool gf = rc == 0;
                            for (double v : g) gf = gf && std::isfinite(v);
                            check(gf, "finite derivative integrals", la, lb, lc, ld);
                        }
                    }
    // orientations the harness must refuse (not evaluate)
    double z[3] = {0, 0, 0}, one = 1.0, out[1296];
    check(chk_eri_quartet(0, 1, 0, 0, z, z, z, z, 1, &one, &one, 1, &one, &one, 1, &one, &one, 1, &one, &one, 0, out) != 0,
          "l_A < l_B refused", 0, 1, 0, 0);
    check(chk_eri_quartet(0, 0, 1, 0, z, z, z, z, 1, &one, &one, 1, &one, &one, 1, &one, &one, 1, &one, &one, 0, out) != 0,
          "class(AB) < class(CD) refused", 0, 0, 1, 0);
    if (fails) {
        std::printf("san_check: %d failures\n", fails);
        return 1;
    }
    std::printf("san_check: ok (%zu rule evaluations, all 21 class combinations x 4 in FP64 and FP32)\n",
                Ts.size() * (sizeof codes / sizeof codes[0]));
    return 0;
}
