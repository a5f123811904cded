# cython: language_level=3, boundscheck=False, wraparound=False
"""Cython binding of the direct J/K engine (C ABI of include/hfx.h), the
counterpart of hfsolver's Cython layer (hfsolver/special.pyx:1-11 over
hfsolver/c_hfsolver.pxd:1-9): thin wrappers that pass numpy buffers to the C
entry points and raise on a non-zero status.

Import it through hfsolver_amd.hfx, which loads torch's HIP runtime first.
Engine has the interface of hfsolver_amd.engine.Engine, so rhf_gauss, doscf
and the autograd Function run on either binding.
"""
cimport c_hfx

import numpy as np


class HfxError(RuntimeError):
    pass


cdef class Stats:
    """Statistics of one call (the fields of hfx_stats)."""
    cdef readonly double seconds, quartets_fp64, quartets_fp32, flops_fp64, flops_fp32, jk_seconds
    cdef readonly double integrals_fp64, integrals_fp32


cdef Stats _stats(const c_hfx.hfx_stats *s):
    cdef Stats r = Stats()
    r.seconds = s.seconds
    r.quartets_fp64 = s.quartets_fp64
    r.quartets_fp32 = s.quartets_fp32
    r.flops_fp64 = s.flops_fp64
    r.flops_fp32 = s.flops_fp32
    r.jk_seconds = s.jk_seconds
    r.integrals_fp64 = s.integrals_fp64
    r.integrals_fp32 = s.integrals_fp32
    return r


cdef inline double* _dp(double[::1] a):
    return &a[0] if a is not None else NULL


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


cdef class Engine:
    """One geometry's direct J/K engine (owns device memory until close())."""
    cdef c_hfx.hfx_t *h
    cdef readonly int n, nat
    cdef public object basis

    def __cinit__(self):
        self.h = NULL

    def __init__(self, basis, double schwarz_threshold=1e-12, double fp32_threshold=0.0, int device=0,
                 verbose=False, int kernel_mode=0):
        cdef c_hfx.hfx_opts o
        c_hfx.hfx_default_opts(&o)
        o.schwarz_threshold = schwarz_threshold
        o.fp32_threshold = fp32_threshold
        o.device = device
        o.verbose = int(verbose)
        o.kernel_mode = kernel_mode
        b = basis
        cdef int[::1] nprim = _i32(b.nprim), istart = _i32(b.istart), l = _i32(b.l), m = _i32(b.m), nn = _i32(b.n)
        cdef int[::1] atno = _i32(b.atno)
        cdef double[::1] x = _f64(b.x), y = _f64(b.y), z = _f64(b.z), coef = _f64(b.coef), alpha = _f64(b.alpha)
        cdef double[::1] xyz = _f64(np.asarray(b.xyz).reshape(-1))
        cdef c_hfx.hfx_t *h = NULL
        rc = c_hfx.hfx_create(b.nbf, &nprim[0], &istart[0], &x[0], &y[0], &z[0], &l[0], &m[0], &nn[0],
                              coef.shape[0], &coef[0], &alpha[0], atno.shape[0], &atno[0], &xyz[0], &o, &h)
        self.h = h
        if rc != 0:
            msg = c_hfx.hfx_last_error(h).decode() if h != NULL else "hfx_create failed"
            self.close()
            raise HfxError(msg)
        self.basis = basis
        self.n = b.nbf
        self.nat = atno.shape[0]

    def __dealloc__(self):
        if self.h != NULL:
            c_hfx.hfx_destroy(self.h)
            self.h = NULL

    def close(self):
        if self.h != NULL:
            c_hfx.hfx_destroy(self.h)
            self.h = NULL

    cdef _check(self, int rc):
        if rc != 0:
            raise HfxError(c_hfx.hfx_last_error(self.h).decode())

    cdef c_hfx.hfx_t* _live(self) except NULL:
        if self.h == NULL:
            raise HfxError("engine is closed")
        return self.h

    def fock_2e(self, P, want_jk=False):
        """G(P) = J - K/2 (and J, K) from host arrays: (G, J, K, Stats)."""
        cdef c_hfx.hfx_t *h = self._live()
        n = self.n
        Pf = np.asfortranarray(P, dtype=np.float64)
        if Pf.shape != (n, n):
            raise ValueError("P must be %d x %d" % (n, n))
        G = np.zeros((n, n), order="F")
        J = np.zeros((n, n), order="F") if want_jk else None
        K = np.zeros((n, n), order="F") if want_jk else None
        cdef double[::1] pv = Pf.reshape(-1, order="F"), gv = G.reshape(-1, order="F")
        cdef double[::1] jv = J.reshape(-1, order="F") if want_jk else None
        cdef double[::1] kv = K.reshape(-1, order="F") if want_jk else None
        cdef c_hfx.hfx_stats st
        self._check(c_hfx.hfx_fock_2e(h, &pv[0], &gv[0], _dp(jv), _dp(kv), &st))
        return G, J, K, _stats(&st)

    def fock_2e_device(self, dP, dG, dJ=None, dK=None):
        """Device form: torch float64 tensors of n*n elements on the engine's GPU.
        torch's stream that produced dP is synchronised first; outputs are complete on return."""
        import torch
        cdef c_hfx.hfx_t *h = self._live()
        torch.cuda.current_stream(dP.device).synchronize()
        cdef size_t p = dP.data_ptr(), g = dG.data_ptr()
        cdef size_t j = 0 if dJ is None else dJ.data_ptr()
        cdef size_t k = 0 if dK is None else dK.data_ptr()
        cdef c_hfx.hfx_stats st
        self._check(c_hfx.hfx_fock_2e_device(h, <const double*>p, <double*>g, <double*>j, <double*>k, &st))
        return _stats(&st)

    def one_electron(self):
        cdef c_hfx.hfx_t *h = self._live()
        n = self.n
        S = np.zeros((n, n), order="F")
        T = np.zeros((n, n), order="F")
        V = np.zeros((n, n), order="F")
        cdef double[::1] s = S.reshape(-1, order="F"), t = T.reshape(-1, order="F"), v = V.reshape(-1, order="F")
        self._check(c_hfx.hfx_one_electron(h, &s[0], &t[0], &v[0]))
        return S, T, V

    def grad_2e(self, P):
        cdef c_hfx.hfx_t *h = self._live()
        cdef double[::1] pv = np.asfortranarray(P, dtype=np.float64).reshape(-1, order="F")
        g = np.zeros(3 * self.nat)
        cdef double[::1] gv = g
        self._check(c_hfx.hfx_grad_2e(h, &pv[0], &gv[0]))
        return g.reshape(self.nat, 3)

    def grad_2e_stats(self):
        cdef c_hfx.hfx_t *h = self._live()
        cdef c_hfx.hfx_stats st
        self._check(c_hfx.hfx_grad_2e_stats(h, &st))
        return _stats(&st)

    def grad_1e(self, P, W):
        cdef c_hfx.hfx_t *h = self._live()
        cdef double[::1] pv = np.asfortranarray(P, dtype=np.float64).reshape(-1, order="F")
        cdef double[::1] wv = np.asfortranarray(W, dtype=np.float64).reshape(-1, order="F")
        g = np.zeros(3 * self.nat)
        cdef double[::1] gv = g
        self._check(c_hfx.hfx_grad_1e(h, &pv[0], &wv[0], &gv[0]))
        return g.reshape(self.nat, 3)

    def set_geometry(self, xyz):
        cdef c_hfx.hfx_t *h = self._live()
        x = _f64(np.asarray(xyz).reshape(self.nat, 3))
        cdef double[::1] xv = x.reshape(-1)
        self._check(c_hfx.hfx_set_geometry(h, &xv[0]))
        self.basis.move(x)

    def eri_shells(self, quartets):
        """Integral blocks of explicit shell quartets [(A,B,C,D), ...] -> list of arrays."""
        cdef c_hfx.hfx_t *h = self._live()
        q = _i32(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
        _, _, Ls, _ = self.shells()
        nc = lambda L: (L + 1) * (L + 2) // 2
        sizes = [nc(Ls[a]) * nc(Ls[b]) * nc(Ls[c]) * nc(Ls[d]) for a, b, c, d in q]
        out = np.zeros(max(1, int(sum(sizes))))
        cdef int[::1] qv = q.reshape(-1)
        cdef double[::1] ov = out
        if len(q):
            self._check(c_hfx.hfx_eri_shells(h, len(q), &qv[0], &ov[0]))
        res, o = [], 0
        for (a, b, c, d), s in zip(q, sizes):
            res.append(out[o:o + s].reshape(nc(Ls[a]), nc(Ls[b]), nc(Ls[c]), nc(Ls[d])))
            o += s
        return res

    def int2(self):
        """Every unique (ij|kl) in the reference's 8-fold packed order (store mode)."""
        cdef c_hfx.hfx_t *h = self._live()
        n = self.n
        mm = n * (n + 1) // 2
        out = np.zeros(mm * (mm + 1) // 2)
        cdef double[::1] ov = out
        self._check(c_hfx.hfx_int2(h, &ov[0], out.size))
        return out

    def shells(self):
        cdef c_hfx.hfx_t *h = self._live()
        cdef int ns = 0
        self._check(c_hfx.hfx_shells(h, &ns, NULL, NULL, NULL))
        f0 = np.zeros(ns, np.int32)
        Ls = np.zeros(ns, np.int32)
        at = np.zeros(ns, np.int32)
        cdef int[::1] fv = f0, lv = Ls, av = at
        self._check(c_hfx.hfx_shells(h, &ns, &fv[0], &lv[0], &av[0]))
        return ns, f0, Ls, at

    def pair_bounds(self):
        cdef c_hfx.hfx_t *h = self._live()
        cdef int npr = 0
        self._check(c_hfx.hfx_pair_bounds(h, &npr, NULL, NULL, NULL))
        A = np.zeros(npr, np.int32)
        B = np.zeros(npr, np.int32)
        Q = np.zeros(npr)
        cdef int[::1] av = A, bv = B
        cdef double[::1] qv = Q
        self._check(c_hfx.hfx_pair_bounds(h, &npr, &av[0], &bv[0], &qv[0]))
        return A, B, Q

    def set_density_screening(self, double threshold):
        cdef c_hfx.hfx_t *h = self._live()
        self._check(c_hfx.hfx_set_density_screening(h, threshold))

    def screening(self):
        cdef c_hfx.hfx_t *h = self._live()
        cdef double a = 0, b = 0, c = 0
        self._check(c_hfx.hfx_screening(h, &a, &b, &c))
        return a, b, c

    def work_model(self):
        """((model FP64, model FP32), (lower bound FP64, lower bound FP32)) FLOPs of one Fock build."""
        cdef c_hfx.hfx_t *h = self._live()
        cdef double m[2]
        cdef double lo[2]
        self._check(c_hfx.hfx_work_model(h, m, lo))
        return (m[0], m[1]), (lo[0], lo[1])
