/*
 * hfx.h - C ABI of the MI355X direct-SCF two-electron engine.
 *
 * Drop-in boundary for hfsolver's closed-shell RHF path. In the reference the
 * two-electron work is split between
 *   - the integral plug-in `getints2` (src/gaussians_hfsolver.f90:13-34), which
 *     calls C `getInts` (src/qc.h:35-39, src/qc.c:95-132) once per SCF to fill
 *     the packed int2 array, and
 *   - `G = formG(P, TT)` (src/scf.f90:26-37, call site src/scf.f90:109),
 *     evaluated every SCF iteration from the stored n^4 tensor TT.
 * This engine replaces both by one direct call per iteration: hfx_fock_2e
 * evaluates the Schwarz-screened electron-repulsion integrals on the GPU and
 * contracts them with the density on the fly (they are never stored).
 *
 * Conventions (same as the reference's C kernels):
 *   - basis description: the per-basis-function arrays gaussints already holds
 *     (src/basis.f90:541-567): nprim[nbf], istart[nbf] (0-BASED, as
 *     gaussints passes istart-1), centre x/y/z[nbf] in bohr, Cartesian powers
 *     l/m/n[nbf] in get_cartesian_shell order (src/basis.f90:371-409),
 *     coef[n2] (normalised, src/basis.f90:486-487) and alpha[n2];
 *   - matrices are n x n, FP64, column-major (S[i + j*nbf], src/qc.c:15),
 *     symmetric, full storage;
 *   - every call is synchronous: outputs are valid on return;
 *   - status: 0 on success, non-zero on error; hfx_last_error() explains.
 * The caller owns every input and output buffer; the handle owns its device
 * memory (basis, shell-pair data, screened pair lists, Rys tables) until
 * hfx_destroy. One handle per host thread; calls on one handle are serial.
 */
#ifndef HFX_H
#define HFX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfx_engine hfx_t;

typedef struct {
    double schwarz_threshold; /* drop quartets with Q_ab * Q_cd below this (default 1e-12) */
    double fp32_threshold;    /* quartets with Q_ab * Q_cd below this use FP32 ERI arithmetic;
                                 0 = all FP64 (default 0) */
    int device;               /* HIP device ordinal (default 0) */
    int verbose;              /* print set-up statistics to stderr (default 0) */
    int kernel_mode;          /* 0: automatic (K rows accumulated in LDS where they fit),
                                 1: per-quartet atomics only (default 0) */
} hfx_opts;

typedef struct {
    double seconds;           /* wall time of the call, device-synchronised */
    double quartets_fp64;     /* shell quartets evaluated in FP64 */
    double quartets_fp32;     /* shell quartets evaluated in FP32 */
    double flops_fp64;        /* algorithmic FLOPs of the FP64 part (frozen per-class model) */
    double flops_fp32;        /* algorithmic FLOPs of the FP32 part */
    double jk_seconds;        /* J/K kernel time, HIP events on the engine's stream */
    double integrals_fp64;    /* contracted Cartesian integrals (ab|cd) of the FP64 quartets */
    double integrals_fp32;    /* same for the FP32 quartets */
} hfx_stats;

/* Fill *o with the defaults. Returns 0. */
int hfx_default_opts(hfx_opts *o);

/*
 * Create an engine for one geometry.
 * Replaces the set-up half of getints2/getInts (src/gaussians_hfsolver.f90:13-34,
 * src/qc.h:35-39): the same per-basis-function arrays, regrouped into shells
 * inside (as src/gaussians_libint.f90:59-122 does) and screened by Schwarz
 * bounds computed on the GPU. nat/atno/xyz (xyz[3*nat], bohr, atom-major as
 * the Fortran xyz(3,nat)) give the nuclei: V, forces and hfx_set_geometry
 * need them; every basis-function centre must coincide with one atom.
 */
int hfx_create(int nbf, const int *nprim, const int *istart, const double *x, const double *y, const double *z,
               const int *l, const int *m, const int *n, int n2, const double *coef, const double *alpha, int nat,
               const int *atno, const double *xyz, const hfx_opts *o, hfx_t **h);

/*
 * Two-electron Fock matrix of a density, host buffers.
 * Replaces G = formG(P, TT) (src/scf.f90:26-37, :109):
 *   J_ij = sum_kl P_kl (ij|kl),  K_ij = sum_kl P_kl (il|kj),  G = J - K/2.
 * P: n x n input (symmetrised as (P + P^T)/2). G, J, K: n x n outputs; any of
 * them may be NULL. stats may be NULL. A non-finite element of P (NaN, Inf)
 * fails the call (status 1, its position in hfx_last_error); the same holds
 * for the P and W of hfx_grad_2e and hfx_grad_1e.
 */
int hfx_fock_2e(hfx_t *h, const double *P, double *G, double *J, double *K, hfx_stats *stats);

/*
 * Same on device buffers (HIP device pointers of n*n doubles, resident in
 * HBM): no host transfers. Used by benchmarks and by callers that keep the
 * density on the GPU. The engine works on streams of its own: the writes that
 * produced dP must be complete when it is called (synchronise the producing
 * stream first). Synchronous: dG, dJ, dK are complete on return. dP is not
 * scanned: a non-finite element gives non-finite elements of dG (the SCF
 * drivers stop on a non-finite energy).
 */
int hfx_fock_2e_device(hfx_t *h, const double *dP, double *dG, double *dJ, double *dK, hfx_stats *stats);

/*
 * One-electron matrices S, T, V (n x n, any may be NULL) for the current
 * geometry. Counterpart of getS/getT/getV (src/qc.c:6-93, src/qc.h:13-33);
 * V includes the nuclear charges atno. Evaluated on the GPU (one_e_kernel).
 */
int hfx_one_electron(hfx_t *h, double *S, double *T, double *V);

/*
 * Two-electron part of the RHF nuclear gradient (no reference counterpart):
 *   g_X = sum_{ijkl} [1/2 P_ij P_kl - 1/4 P_ik P_jl] d(ij|kl)/dX,
 * evaluated with derivative integrals on the GPU. g: 3*nat (x,y,z per atom).
 */
int hfx_grad_2e(hfx_t *h, const double *P, double *g);

/*
 * Statistics of the last hfx_grad_2e call on h (zeros before the first):
 * seconds (wall time), the shell quartets and contracted integrals it
 * differentiated (FP64 / FP32 launches), their FLOPs by the frozen operation
 * count of the derivative kernel (flops_fp64 / flops_fp32), and in jk_seconds
 * the derivative kernels' time by HIP events on the engine's stream.
 */
int hfx_grad_2e_stats(const hfx_t *h, hfx_stats *st);

/*
 * One-electron part of the RHF gradient:
 *   g_X = sum_ij P_ij dH_ij/dX - sum_ij W_ij dS_ij/dX  (Hellmann-Feynman term included),
 * W = energy-weighted density. g: 3*nat. Evaluated on the GPU (one_e_kernel).
 */
int hfx_grad_1e(hfx_t *h, const double *P, const double *W, double *g);

/* Move the nuclei (xyz[3*nat], bohr); basis functions follow their atom.
 * Rebuilds pair data, Schwarz bounds and quartet lists. */
int hfx_set_geometry(hfx_t *h, const double *xyz);

/*
 * Integrals of explicit shell quartets (testing and store-mode use).
 * shells[4*k .. 4*k+3] = (A, B, C, D) shell indices (shells numbered in
 * basis-function order); out receives the blocks one after another, each as
 * out[((a*nB + b)*nC + c)*nD + d] with a, b, c, d the Cartesian components of
 * A, B, C, D in reference order (d fastest).
 * Evaluated on the GPU in FP64 with no screening.
 */
int hfx_eri_shells(hfx_t *h, int nq, const int *shells, double *out);

/*
 * hfx_eri_shells without any memory allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, kernels and results as hfx_eri_shells, but the quartet plans and
 * the integral blocks are staged in a fixed device scratch array of 16 MiB per
 * GPU (allocated with the library's code object, held until the process ends)
 * instead of device memory allocated per call; no hipMalloc, hipFree or host
 * heap allocation happens. A request whose plans and integrals do not fit
 * (about 20 + 8 x integrals bytes per quartet, e.g. 1 600 (dd|dd) quartets)
 * fails with an error naming the sizes. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (from several host threads or
 * handles) race and must be serialised by the caller.
 */
int hfx_eri_shells_noalloc(hfx_t *h, int nq, const int *shells, double *out);

/*
 * Store mode: every unique two-electron integral over the basis functions,
 * written into the 8-fold packed array int2[size], size = m(m+1)/2 with
 * m = nbf(nbf+1)/2, at the 0-based index of src/cints.c:327-347
 * (ij = i(i+1)/2 + j for i >= j, the same for kl, then the larger pair index
 * first; 64-bit here).
 * Replaces getints2 (src/gaussians_hfsolver.f90:13-34) -> getInts
 * (src/qc.c:95-132, src/qc.h:35-39), which fill the same array on the CPU.
 * Evaluated on the GPU in FP64; no Schwarz screening (shell pairs whose every
 * primitive pair was screened out at hfx_create give zeros: create the handle
 * with schwarz_threshold = 0 to keep every primitive pair). The array is
 * computed in device memory (size * 8 bytes) and copied to int2.
 */
int hfx_int2(hfx_t *h, double *int2, long long size);

/* Shell structure: number of shells; per shell its first basis function,
 * angular momentum and atom (arrays of length nshell, may be NULL). */
int hfx_shells(const hfx_t *h, int *nshell, int *first_bf, int *L, int *atom);

/* All shell pairs of the current geometry with their Schwarz bounds
 * Q_AB = sqrt(max |(ab|ab)|) computed on the GPU: *npairs on return; A, B, Q
 * (arrays of nshell*(nshell+1)/2, may be NULL) receive the shells and bounds. */
int hfx_pair_bounds(const hfx_t *h, int *npairs, int *A, int *B, double *Q);

/*
 * Orbital integrals from the packed AO integrals of hfx_int2: the four-index
 * transformation of transform_int2 (src/mbpt.f90:11-79), on the GPU:
 *   moint2[index(i,j,k,l)] = sum C(mu,i) C(nu,j) C(al,k) C(be,l) int2[index(mu,nu,al,be)],
 * C: n x n column-major (column i = orbital i), int2 / moint2: the packed
 * 8-fold vectors of `size` = m(m+1)/2 elements, m = n(n+1)/2. Needs
 * 2 n^4 doubles of device memory (checked; an error names the shortfall).
 */
int hfx_transform_int2(hfx_t *h, const double *int2, const double *C, double *moint2, long long size);

/* Density-weighted screening of the following Fock builds (no counterpart in
 * the reference, whose formG contracts every stored integral; used by
 * incremental builds G(P) = G(P_prev) + G(P - P_prev) in the SCF): a screened
 * quartet is skipped when Q_ab Q_cd times the largest |D| over its six density
 * blocks (ab, cd, ac, ad, bc, bd; D the symmetrised density, per shell-pair
 * block) is below threshold. 0 (the default) turns it off. Applies to the
 * row-mode launches (all classes at the benchmark sizes). */
int hfx_set_density_screening(hfx_t *h, double threshold);

/* Statistics of the last set-up: number of significant shell pairs and of
 * screened shell quartets in FP64 / FP32. */
int hfx_screening(const hfx_t *h, double *npairs, double *nq_fp64, double *nq_fp32);

/* Work of one Fock build at the current set-up (no counterpart in the
 * reference): model[0], model[1] = modelled FLOPs of the FP64 / FP32 launches
 * (the frozen per-class operation counts hfx_stats reports); lower[0],
 * lower[1] = the algorithm-neutral lower bound, the sum over screened quartets
 * of K_q * 3 nroots(L_q) * N_q (primitive quartets K_q times the Rys
 * assembly of each of the N_q integrals: nroots = floor(L_q / 2) + 1 roots,
 * a product and a multiply-add each). Either array may be NULL. */
int hfx_work_model(const hfx_t *h, double *model, double *lower);

const char *hfx_last_error(const hfx_t *h);
void hfx_destroy(hfx_t *h);

#ifdef __cplusplus
}
#endif

#endif /* HFX_H */
