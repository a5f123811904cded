# Declarations of the engine's C ABI (include/hfx.h) for Cython, in the style
# of hfsolver's c_hfsolver.pxd (hfsolver/c_hfsolver.pxd:1-9): the C signatures
# as they are, pointers for arrays.

cdef extern from "hfx.h":

    ctypedef struct hfx_t:
        pass

    ctypedef struct hfx_opts:
        double schwarz_threshold
        double fp32_threshold
        int device
        int verbose
        int kernel_mode

    ctypedef struct hfx_stats:
        double seconds
        double quartets_fp64
        double quartets_fp32
        double flops_fp64
        double flops_fp32
        double jk_seconds
        double integrals_fp64
        double integrals_fp32

    int hfx_default_opts(hfx_opts *o)
    int hfx_create(int nbf, const int *nprim, const int *istart, const double *x, const double *y,
                   const double *z, const int *l, const int *m, const int *n, int n2, const double *coef,
                   const double *alpha, int nat, const int *atno, const double *xyz, const hfx_opts *o,
                   hfx_t **h)
    int hfx_fock_2e(hfx_t *h, const double *P, double *G, double *J, double *K, hfx_stats *stats)
    int hfx_fock_2e_device(hfx_t *h, const double *dP, double *dG, double *dJ, double *dK, hfx_stats *stats)
    int hfx_one_electron(hfx_t *h, double *S, double *T, double *V)
    int hfx_grad_2e(hfx_t *h, const double *P, double *g)
    int hfx_grad_2e_stats(const hfx_t *h, hfx_stats *st)
    int hfx_grad_1e(hfx_t *h, const double *P, const double *W, double *g)
    int hfx_set_geometry(hfx_t *h, const double *xyz)
    int hfx_eri_shells(hfx_t *h, int nq, const int *shells, double *out)
    int hfx_eri_shells_noalloc(hfx_t *h, int nq, const int *shells, double *out)
    int hfx_int2(hfx_t *h, double *int2, long long size)
    int hfx_transform_int2(hfx_t *h, const double *int2, const double *C, double *moint2, long long size)
    int hfx_shells(const hfx_t *h, int *nshell, int *first_bf, int *L, int *atom)
    int hfx_pair_bounds(const hfx_t *h, int *npairs, int *A, int *B, double *Q)
    int hfx_set_density_screening(hfx_t *h, double threshold)
    int hfx_screening(const hfx_t *h, double *npairs, double *nq_fp64, double *nq_fp32)
    int hfx_work_model(const hfx_t *h, double *model, double *lower)
    const char *hfx_last_error(const hfx_t *h)
    void hfx_destroy(hfx_t *h)
