"""ctypes binding of the native engine (libhfx.so, C ABI in include/hfx.h).

The library is built in-tree by ``python -m hfsolver_amd.build``. There is no
fallback: if it is missing, or no HIP device is present, every entry point
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhfx.so")

_lib = None


class HfxError(RuntimeError):
    pass


class HfxOpts(ctypes.Structure):
    _fields_ = [("schwarz_threshold", ctypes.c_double), ("fp32_threshold", ctypes.c_double),
                ("device", ctypes.c_int), ("verbose", ctypes.c_int), ("kernel_mode", ctypes.c_int)]


class HfxStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("quartets_fp64", ctypes.c_double),
                ("quartets_fp32", ctypes.c_double), ("flops_fp64", ctypes.c_double),
                ("flops_fp32", ctypes.c_double), ("jk_seconds", ctypes.c_double),
                ("integrals_fp64", ctypes.c_double), ("integrals_fp32", ctypes.c_double)]


# name -> (restype, argtypes); every symbol declared in include/hfx.h
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
SIGNATURES = {
    "hfx_default_opts": (_I, [ctypes.POINTER(HfxOpts)]),
    "hfx_create": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _I, _P, _P, ctypes.POINTER(HfxOpts),
                        ctypes.POINTER(_P)]),
    "hfx_fock_2e": (_I, [_P, _P, _P, _P, _P, ctypes.POINTER(HfxStats)]),
    "hfx_fock_2e_device": (_I, [_P, _P, _P, _P, _P, ctypes.POINTER(HfxStats)]),
    "hfx_one_electron": (_I, [_P, _P, _P, _P]),
    "hfx_grad_2e": (_I, [_P, _P, _P]),
    "hfx_grad_2e_stats": (_I, [_P, ctypes.POINTER(HfxStats)]),
    "hfx_grad_1e": (_I, [_P, _P, _P, _P]),
    "hfx_set_geometry": (_I, [_P, _P]),
    "hfx_eri_shells": (_I, [_P, _I, _P, _P]),
    "hfx_eri_shells_noalloc": (_I, [_P, _I, _P, _P]),
    "hfx_int2": (_I, [_P, _P, ctypes.c_longlong]),
    "hfx_transform_int2": (_I, [_P, _P, _P, _P, ctypes.c_longlong]),
    "hfx_shells": (_I, [_P, ctypes.POINTER(_I), _P, _P, _P]),
    "hfx_screening": (_I, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "hfx_set_density_screening": (_I, [_P, _D]),
    "hfx_pair_bounds": (_I, [_P, ctypes.POINTER(_I), _P, _P, _P]),
    "hfx_work_model": (_I, [_P, _P, _P]),
    "hfx_last_error": (ctypes.c_char_p, [_P]),
    "hfx_destroy": (None, [_P]),
}


def lib():
    """Load libhfx.so (raises HfxError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HfxError("native library %s is not built (python -m hfsolver_amd.build)" % LIB_PATH)
        # torch for ROCm ships its own HIP and HSA runtimes (same sonames as
        # /opt/rocm's). Loaded first, they also serve libhfx.so, so that the
        # process has one runtime and torch device buffers can be handed to the
        # engine; loaded after libhfx.so they would start a second HSA runtime,
        # which finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Engine:
    """One geometry's direct J/K engine (owns device memory until close())."""

    def __init__(self, basis, schwarz_threshold=1e-12, fp32_threshold=0.0, device=0, verbose=False, kernel_mode=0):
        L = lib()
        self.basis = basis
        opts = HfxOpts()
        L.hfx_default_opts(ctypes.byref(opts))
        opts.schwarz_threshold = schwarz_threshold
        opts.fp32_threshold = fp32_threshold
        opts.device = device
        opts.verbose = int(verbose)
        opts.kernel_mode = kernel_mode
        h = ctypes.c_void_p()
        b = basis
        self._keep = [b.nprim, b.istart, b.x, b.y, b.z, b.l, b.m, b.n, b.coef, b.alpha, b.atno, b.xyz]
        rc = L.hfx_create(b.nbf, _ptr(b.nprim), _ptr(b.istart), _ptr(b.x), _ptr(b.y), _ptr(b.z), _ptr(b.l),
                          _ptr(b.m), _ptr(b.n), len(b.coef), _ptr(b.coef), _ptr(b.alpha), len(b.atno),
                          _ptr(b.atno), _ptr(b.xyz), ctypes.byref(opts), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = L.hfx_last_error(h).decode() if h.value else "hfx_create failed"
            self.close()
            raise HfxError(msg)
        self.n = b.nbf
        self.nat = len(b.atno)

    def _check(self, rc):
        if rc != 0:
            raise HfxError(lib().hfx_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().hfx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fock_2e(self, P, want_jk=False):
        """G(P) = J - K/2 (and J, K) from host arrays; returns (G, J, K, stats)."""
        n = self.n
        P = np.asfortranarray(P, dtype=np.float64)
        if P.shape != (n, n):
            raise ValueError("P must be %d x %d" % (n, n))
        G = np.zeros((n, n), order="F")
        J = np.zeros((n, n), order="F") if want_jk else None
        K = np.zeros((n, n), order="F") if want_jk else None
        st = HfxStats()
        self._check(lib().hfx_fock_2e(self._h, _ptr(P), _ptr(G), _ptr(J), _ptr(K), ctypes.byref(st)))
        return G, J, K, st

    def fock_2e_device(self, dP, dG, dJ=None, dK=None):
        """Device-pointer form: dP/dG/dJ/dK are torch CUDA (HIP) float64 tensors of n*n elements.

        The engine runs on streams of its own, so the torch stream that produced
        dP is synchronised first; the outputs are complete on return."""
        import torch
        torch.cuda.current_stream(dP.device).synchronize()
        st = HfxStats()
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        self._check(lib().hfx_fock_2e_device(self._h, ptr(dP), ptr(dG), ptr(dJ), ptr(dK), ctypes.byref(st)))
        return st

    def one_electron(self):
        n = self.n
        S = np.zeros((n, n), order="F")
        T = np.zeros((n, n), order="F")
        V = np.zeros((n, n), order="F")
        self._check(lib().hfx_one_electron(self._h, _ptr(S), _ptr(T), _ptr(V)))
        return S, T, V

    def grad_2e(self, P):
        P = np.asfortranarray(P, dtype=np.float64)
        g = np.zeros(3 * self.nat)
        self._check(lib().hfx_grad_2e(self._h, _ptr(P), _ptr(g)))
        return g.reshape(self.nat, 3)

    def grad_2e_stats(self):
        """Statistics of the last grad_2e call (HfxStats: time, quartets, model FLOPs, kernel seconds)."""
        st = HfxStats()
        self._check(lib().hfx_grad_2e_stats(self._h, ctypes.byref(st)))
        return st

    def grad_1e(self, P, W):
        P = np.asfortranarray(P, dtype=np.float64)
        W = np.asfortranarray(W, dtype=np.float64)
        g = np.zeros(3 * self.nat)
        self._check(lib().hfx_grad_1e(self._h, _ptr(P), _ptr(W), _ptr(g)))
        return g.reshape(self.nat, 3)

    def set_geometry(self, xyz):
        xyz = _f64(np.asarray(xyz).reshape(self.nat, 3))
        self._check(lib().hfx_set_geometry(self._h, _ptr(xyz)))
        self.basis.move(xyz)

    def eri_shells(self, quartets, noalloc=False):
        """Integral blocks of explicit shell quartets [(A,B,C,D), ...] -> list of arrays.
        noalloc: through hfx_eri_shells_noalloc (fixed device scratch, no allocation)."""
        q = np.ascontiguousarray(np.asarray(quartets, dtype=np.int32).reshape(-1, 4))
        _, _, Ls, _ = self.shells()
        nc = lambda l: (l + 1) * (l + 2) // 2
        sizes = [nc(Ls[a]) * nc(Ls[b]) * nc(Ls[c]) * nc(Ls[d]) for a, b, c, d in q]
        out = np.zeros(int(sum(sizes)))
        fn = lib().hfx_eri_shells_noalloc if noalloc else lib().hfx_eri_shells
        self._check(fn(self._h, len(q), _ptr(q), _ptr(out)))
        res, o = [], 0
        for (a, b, c, d), s in zip(q, sizes):
            res.append(out[o:o + s].reshape(nc(Ls[a]), nc(Ls[b]), nc(Ls[c]), nc(Ls[d])))
            o += s
        return res

    def int2(self):
        """Every unique (ij|kl) in the reference's 8-fold packed order (store mode)."""
        n = self.basis.nbf
        m = n * (n + 1) // 2
        out = np.zeros(m * (m + 1) // 2)
        self._check(lib().hfx_int2(self._h, _ptr(out), out.size))
        return out

    def transform_int2(self, int2, C):
        """Orbital integrals (i j|k l) in the packed order of int2 (transform_int2
        of mbpt.f90, four GPU quarter transforms); C[:, i] = orbital i."""
        int2 = _f64(int2)
        C = np.asfortranarray(C, dtype=np.float64)
        out = np.zeros_like(int2)
        self._check(lib().hfx_transform_int2(self._h, _ptr(int2), _ptr(C), _ptr(out), out.size))
        return out

    def shells(self):
        ns = ctypes.c_int()
        self._check(lib().hfx_shells(self._h, ctypes.byref(ns), None, None, None))
        f0 = np.zeros(ns.value, np.int32)
        Ls = np.zeros(ns.value, np.int32)
        at = np.zeros(ns.value, np.int32)
        self._check(lib().hfx_shells(self._h, ctypes.byref(ns), _ptr(f0), _ptr(Ls), _ptr(at)))
        return ns.value, f0, Ls, at

    def pair_bounds(self):
        """(A, B, Q) of every shell pair (Schwarz bounds from the GPU set-up)."""
        npr = ctypes.c_int()
        self._check(lib().hfx_pair_bounds(self._h, ctypes.byref(npr), None, None, None))
        A = np.zeros(npr.value, np.int32)
        B = np.zeros(npr.value, np.int32)
        Q = np.zeros(npr.value)
        self._check(lib().hfx_pair_bounds(self._h, ctypes.byref(npr), _ptr(A), _ptr(B), _ptr(Q)))
        return A, B, Q

    def work_model(self):
        """((model FP64, model FP32), (lower bound FP64, lower bound FP32)) FLOPs of one Fock build:
        the frozen per-class model and the Rys-assembly lower bound sum_q K_q 3 nroots N_q."""
        m = np.zeros(2)
        lo = np.zeros(2)
        self._check(lib().hfx_work_model(self._h, _ptr(m), _ptr(lo)))
        return (float(m[0]), float(m[1])), (float(lo[0]), float(lo[1]))

    def set_density_screening(self, threshold):
        """Skip quartets whose Schwarz bound times the largest density factor is
        below threshold in the following Fock builds (0: off)."""
        self._check(lib().hfx_set_density_screening(self._h, float(threshold)))

    def screening(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().hfx_screening(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value
