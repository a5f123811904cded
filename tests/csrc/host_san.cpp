// The engine's host runtime (hfsolver_amd/csrc/hfx_host.cpp: shell regrouping,
// primitive-pair screening, Schwarz-sorted class lists, ket tables, row-mode
// block lists, buffers of every entry point) compiled with AddressSanitizer and
// UndefinedBehaviorSanitizer for the host, linked with the normal gfx950 kernel
// objects, and driven through every entry point of include/hfx.h on a real
// basis. An invalid host access or undefined behaviour stops the program with a
// sanitizer report; a failed call stops it with the engine's message.
//
//   host_san BASIS_FILE OUT_FILE
//
// BASIS_FILE (little-endian, written by tests/test_gpu_host_sanitizers.py):
//   int32 nbf, n2, nat; int32 nprim[nbf], istart[nbf]; f64 x[nbf], y[nbf],
//   z[nbf]; int32 l[nbf], m[nbf], n[nbf]; f64 coef[n2], alpha[n2];
//   int32 atno[nat]; f64 xyz[3 nat]; f64 P[nbf nbf]
// OUT_FILE: f64 G[nbf nbf] of the mixed-precision engine at the input
// geometry, then f64 g[3 nat] of its two-electron gradient.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hfx.h"

namespace {

template <typename T>
void rd(FILE* f, T* p, size_t n) {
    if (n && std::fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "host_san: short basis file\n");
        std::exit(2);
    }
}

void ok(hfx_t* h, int rc, const char* what) {
    if (rc != 0) {
        std::fprintf(stderr, "host_san: %s failed: %s\n", what, h ? hfx_last_error(h) : "(no handle)");
        std::exit(1);
    }
    std::printf("host_san: %s ok\n", what);
    std::fflush(stdout);
}

void expect_failure(hfx_t* h, int rc, const char* what) {
    if (rc == 0) {
        std::fprintf(stderr, "host_san: %s was accepted\n", what);
        std::exit(1);
    }
    std::printf("host_san: %s rejected (%s)\n", what, h ? hfx_last_error(h) : "no handle");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: host_san BASIS_FILE OUT_FILE\n");
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "host_san: cannot open %s\n", argv[1]);
        return 2;
    }
    int hdr[3];
    rd(f, hdr, 3);
    const int nbf = hdr[0], n2 = hdr[1], nat = hdr[2];
    std::vector<int> nprim(nbf), istart(nbf), l(nbf), m(nbf), n(nbf), atno(nat);
    std::vector<double> x(nbf), y(nbf), z(nbf), coef(n2), alpha(n2), xyz(3 * (size_t)nat), P((size_t)nbf * nbf);
    rd(f, nprim.data(), nbf);
    rd(f, istart.data(), nbf);
    rd(f, x.data(), nbf);
    rd(f, y.data(), nbf);
    rd(f, z.data(), nbf);
    rd(f, l.data(), nbf);
    rd(f, m.data(), nbf);
    rd(f, n.data(), nbf);
    rd(f, coef.data(), n2);
    rd(f, alpha.data(), n2);
    rd(f, atno.data(), nat);
    rd(f, xyz.data(), 3 * (size_t)nat);
    rd(f, P.data(), (size_t)nbf * nbf);
    std::fclose(f);
    const size_t nn = (size_t)nbf * nbf;

    // mixed-precision engine: FP64 and FP32 quartet lists, row-mode blocks
    hfx_opts o;
    hfx_default_opts(&o);
    o.schwarz_threshold = 1e-12;
    o.fp32_threshold = 1e-5;
    hfx_t* h = nullptr;
    ok(h, hfx_create(nbf, nprim.data(), istart.data(), x.data(), y.data(), z.data(), l.data(), m.data(), n.data(), n2,
                     coef.data(), alpha.data(), nat, atno.data(), xyz.data(), &o, &h),
       "hfx_create (mixed)");
    std::vector<double> S(nn), T(nn), V(nn), G(nn), J(nn), K(nn), g(3 * (size_t)nat), g1(3 * (size_t)nat);
    ok(h, hfx_one_electron(h, S.data(), T.data(), V.data()), "hfx_one_electron");
    hfx_stats st;
    ok(h, hfx_fock_2e(h, P.data(), G.data(), J.data(), K.data(), &st), "hfx_fock_2e");
    std::vector<double> G0 = G;
    ok(h, hfx_grad_2e(h, P.data(), g.data()), "hfx_grad_2e");
    std::vector<double> g0 = g;
    ok(h, hfx_grad_2e_stats(h, &st), "hfx_grad_2e_stats");
    ok(h, hfx_grad_1e(h, P.data(), P.data(), g1.data()), "hfx_grad_1e");
    ok(h, hfx_set_density_screening(h, 1e-11), "hfx_set_density_screening");
    ok(h, hfx_fock_2e(h, P.data(), G.data(), nullptr, nullptr, nullptr), "hfx_fock_2e (density-screened)");
    ok(h, hfx_set_density_screening(h, 0.0), "hfx_set_density_screening (off)");
    double dq[3];
    ok(h, hfx_screening(h, dq, dq + 1, dq + 2), "hfx_screening");
    int nsh = 0;
    ok(h, hfx_shells(h, &nsh, nullptr, nullptr, nullptr), "hfx_shells (count)");
    std::vector<int> fbf(nsh), Ls(nsh), atom(nsh);
    ok(h, hfx_shells(h, &nsh, fbf.data(), Ls.data(), atom.data()), "hfx_shells");
    int npairs = 0;
    ok(h, hfx_pair_bounds(h, &npairs, nullptr, nullptr, nullptr), "hfx_pair_bounds (count)");
    std::vector<int> pa(npairs), pb(npairs);
    std::vector<double> pq(npairs);
    ok(h, hfx_pair_bounds(h, &npairs, pa.data(), pb.data(), pq.data()), "hfx_pair_bounds");
    // explicit quartets over the first and last shells (all angular momenta)
    std::vector<int> q;
    for (int i = 0; i < nsh && (int)q.size() < 4 * 64; i += std::max(1, nsh / 8))
        for (int j = nsh - 1; j >= 0 && (int)q.size() < 4 * 64; j -= std::max(1, nsh / 8)) {
            const int s4[4] = {i, j, j, i};
            q.insert(q.end(), s4, s4 + 4);
        }
    const int nq = (int)q.size() / 4;
    size_t nint = 0;
    for (int k = 0; k < nq; k++) {
        size_t c = 1;
        for (int t = 0; t < 4; t++) c *= (size_t)(Ls[q[4 * k + t]] + 1) * (Ls[q[4 * k + t]] + 2) / 2;
        nint += c;
    }
    std::vector<double> e1(nint), e2(nint);
    ok(h, hfx_eri_shells(h, nq, q.data(), e1.data()), "hfx_eri_shells");
    ok(h, hfx_eri_shells_noalloc(h, nq, q.data(), e2.data()), "hfx_eri_shells_noalloc");
    if (std::memcmp(e1.data(), e2.data(), nint * sizeof(double)) != 0) {
        std::fprintf(stderr, "host_san: hfx_eri_shells_noalloc differs from hfx_eri_shells\n");
        return 1;
    }
    if (nbf <= 100) {
        const long long mp = (long long)nbf * (nbf + 1) / 2;
        std::vector<double> int2((size_t)(mp * (mp + 1) / 2));
        ok(h, hfx_int2(h, int2.data(), (long long)int2.size()), "hfx_int2");
        expect_failure(h, hfx_int2(h, int2.data(), (long long)int2.size() - 1), "hfx_int2 with a short array");
    }
    // rejected input: a NaN in P
    std::vector<double> Pbad = P;
    Pbad[nn / 2] = std::nan("");
    expect_failure(h, hfx_fock_2e(h, Pbad.data(), G.data(), nullptr, nullptr, nullptr), "hfx_fock_2e with NaN in P");
    expect_failure(h, hfx_grad_2e(h, Pbad.data(), g.data()), "hfx_grad_2e with NaN in P");
    // moved nuclei: pair data, bounds and lists rebuilt
    std::vector<double> xyz2 = xyz;
    for (size_t i = 0; i < xyz2.size(); i++) xyz2[i] += 0.01 * std::sin(1.0 + (double)i);
    ok(h, hfx_set_geometry(h, xyz2.data()), "hfx_set_geometry");
    ok(h, hfx_fock_2e(h, P.data(), G.data(), nullptr, nullptr, nullptr), "hfx_fock_2e (moved nuclei)");
    ok(h, hfx_grad_2e(h, P.data(), g.data()), "hfx_grad_2e (moved nuclei)");
    ok(h, hfx_set_geometry(h, xyz.data()), "hfx_set_geometry (back)");
    ok(h, hfx_fock_2e(h, P.data(), G.data(), nullptr, nullptr, nullptr), "hfx_fock_2e (back)");
    hfx_destroy(h);

    // FP64 engine in tile mode (the fallback kernels and their per-sublist ket lists)
    hfx_default_opts(&o);
    o.kernel_mode = 1;
    hfx_t* h2 = nullptr;
    ok(h2, hfx_create(nbf, nprim.data(), istart.data(), x.data(), y.data(), z.data(), l.data(), m.data(), n.data(), n2,
                      coef.data(), alpha.data(), nat, atno.data(), xyz.data(), &o, &h2),
       "hfx_create (tile mode)");
    ok(h2, hfx_fock_2e(h2, P.data(), G.data(), nullptr, nullptr, nullptr), "hfx_fock_2e (tile mode)");
    ok(h2, hfx_grad_2e(h2, P.data(), g.data()), "hfx_grad_2e (tile mode)");
    hfx_destroy(h2);

    // inconsistent basis arrays are rejected before any device work
    std::vector<int> istart_bad = istart;
    istart_bad[nbf - 1] = n2 + 5;
    hfx_t* h3 = nullptr;
    const int rc3 = hfx_create(nbf, nprim.data(), istart_bad.data(), x.data(), y.data(), z.data(), l.data(), m.data(),
                               n.data(), n2, coef.data(), alpha.data(), nat, atno.data(), xyz.data(), &o, &h3);
    expect_failure(h3, rc3, "hfx_create with istart beyond the coefficient array");
    hfx_destroy(h3);

    FILE* out = std::fopen(argv[2], "wb");
    if (!out || std::fwrite(G0.data(), sizeof(double), nn, out) != nn ||
        std::fwrite(g0.data(), sizeof(double), g0.size(), out) != g0.size()) {
        std::fprintf(stderr, "host_san: cannot write %s\n", argv[2]);
        return 2;
    }
    std::fclose(out);
    std::printf("host_san: ok\n");
    return 0;
}
