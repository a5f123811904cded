"""ORACLE - test infrastructure only.

CPU restatement of hfsolver's closed-shell RHF path, used as the checker for
the HIP engine. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product never does.

Python half (this file), each function citing what it restates (paths under
hfsolver's src/):
  read_basis        basis.f90:143-272   (6-31G** text file reader)
  get_basis         basis.f90:275-344
  cartesian_shell   basis.f90:371-409   (Cartesian (l,m,n) order)
  norm_prim         basis.f90:93-98
  norm_contr        basis.f90:100-132
  contracted_gto    basis.f90:425-488   (shells2contracted_gto + fix_normalization)
  gaussints         basis.f90:490-569   (S, T, V, optional packed int2)
  doscf             scf.f90:72-136      (P=0 start, formP, E0, convergence, mixing)
  rhf_gauss         drivers.f90:12-51
  libint_shell_walk gaussians_libint.f90:70-124 (shells from per-function powers)
  transform_int2    mbpt.f90:11-79      (AO -> MO packed integrals)
  mbpt2             mbpt.f90:154-159 + mbpt2.f90:8-31 (closed-shell second order)
  jk_screened       scf.f90:26-37 semantics over Schwarz-screened unique shell
                    quartets (no reference counterpart for the screening)
  run_reference     runs the compiled reference itself (oracle/_ref/ref_driver)
The C half (hfx_oracle.c, loaded through ctypes) holds the Boys function and
the THO integrals.

Parity pins: tests/test_oracle.py checks every function above against the
reference's compiled outputs (tests/golden/*.json), the reference's literature
energies, closed forms (Boys F0, S_ii = 1) and brute-force sums;
transform_int2 and mbpt2 against the reference's CH4 second-order energy
(test_gaussians_molecules.f90:114), the SCF energy rebuilt from the orbital
integrals, and the identity transform.
"""
import ctypes
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
BASIS_FILE = os.path.join(os.path.dirname(HERE), "hfsolver_amd", "data", "p631ss.txt")
BOHR2ANG = 0.529177249          # constants.f90:43
ANG2BOHR = 1.0 / BOHR2ANG       # constants.f90:44

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.orc_boys.argtypes = [i, d, p]
        L.orc_boys.restype = i
        for name in ("orc_getS", "orc_getT", "orc_getInts"):
            getattr(L, name).argtypes = [i] + [p] * 10 + [p]
            getattr(L, name).restype = None
        L.orc_getV.argtypes = [i] + [p] * 10 + [i, p, p, p, p, p]
        L.orc_getV.restype = None
        L.orc_jk.argtypes = [i] + [p] * 10 + [p, p, p, i]
        L.orc_jk.restype = None
        L.orc_eri_bf.argtypes = [i] + [p] * 10 + [i, i, i, i]
        L.orc_eri_bf.restype = d
        L.orc_eri_prim.argtypes = ([d, d, d, i, i, i, d] * 4)
        L.orc_eri_prim.restype = d
        L.orc_grad_2e.argtypes = [i] + [p] * 10 + [p, i, p, p, i]
        L.orc_grad_2e.restype = None
        L.orc_grad_1e.argtypes = [i] + [p] * 10 + [p, i, p, p, p, p, p, p, p]
        L.orc_grad_1e.restype = None
        L.orc_deri_bf.argtypes = [i] + [p] * 10 + [i, i, i, i, i, p]
        L.orc_deri_bf.restype = None
        L.orc_eri_quartets.argtypes = [i] + [p] * 10 + [i, p, p, i]
        L.orc_eri_quartets.restype = d
        L.orc_schwarz.argtypes = [i] + [p] * 10 + [i, p, p, p]
        L.orc_schwarz.restype = None
        L.orc_jk_screened.argtypes = [i] + [p] * 10 + [i, p, p, p, d, p, p, p, i]
        L.orc_jk_screened.restype = d
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def boys(maxm, t):
    F = np.zeros(maxm + 1)
    if lib().orc_boys(maxm, float(t), _ptr(F)) != 0:
        raise ValueError("Fm: only works for 0 <= m <= 500 and t >= 0")
    return F


# ------------------------------------------------------------------ basis ---
_LCHAR = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4, "h": 5, "i": 6}


def read_basis(Z, path=BASIS_FILE):
    """Shells (L, nprim, c, zeta) of atom Z from the 6-31G** file (basis.f90:143-249)."""
    with open(path) as f:
        toks = [line.split() for line in f if line.strip()]
    pos = 0
    while pos < len(toks):
        ZZ = int(toks[pos][0])
        r = int(toks[pos + 1][0])
        pos += 2
        L, nprim, c, zeta = [], [], [], []
        for _ in range(r):
            L.append(_LCHAR[toks[pos][0].lower()])
            k = int(toks[pos][1])
            nprim.append(k)
            pos += 1
            for _ in range(k):
                zeta.append(float(toks[pos][0]))
                c.append(float(toks[pos][1]))
                pos += 1
        if ZZ == Z:
            return L, nprim, c, zeta
    raise ValueError("read_basis: atom Z=%d not found" % Z)


def cartesian_shell(L):
    """(l, m, n) triplets of a Cartesian shell in hfsolver's order (basis.f90:371-409)."""
    out = []
    for k in range(L + 1):
        for n in range(k + 1):
            out.append((L - k, k - n, n))
    return out


def _fact2(n):
    return 1 if n <= 1 else n * _fact2(n - 2)


def _pow2_single(e):
    """2**e evaluated in single precision, then widened: basis.f90:96 writes
    2**(2*(l+m+n)+1.5) with a default-real (REAL(4)) literal 1.5, so the
    reference computes this factor in FP32."""
    return float(np.float32(2.0) ** np.float32(e))


def norm_prim(l, m, n, alpha):
    """basis.f90:93-98 (including the FP32 factor 2**(2L+1.5))"""
    return math.sqrt(_pow2_single(2 * (l + m + n) + 1.5) * alpha ** (l + m + n + 1.5) /
                     _fact2(2 * l - 1) / _fact2(2 * m - 1) / _fact2(2 * n - 1) / math.pi ** 1.5)


def norm_contr(l, m, n, alpha, coefu):
    """basis.f90:100-132"""
    L = l + m + n
    S = 0.0
    for i in range(len(alpha)):
        for j in range(len(alpha)):
            S += coefu[i] * coefu[j] / (alpha[i] + alpha[j]) ** (L + 1.5)
    return 1 / math.pi ** 0.75 * math.sqrt(2 ** L / _fact2(2 * l - 1) / _fact2(2 * m - 1) / _fact2(2 * n - 1)) \
        / math.sqrt(S)


class Basis:
    """Per-basis-function SoA arrays as gaussints holds them (basis.f90:541-550)."""

    def __init__(self, atno, xyz):
        atno = [int(z) for z in atno]
        xyz = np.asarray(xyz, dtype=float).reshape(-1, 3)
        nprim, istart, cen, powr, coef, alpha, bf_atom = [], [], [], [], [], [], []
        for a, Z in enumerate(atno):
            Ls, nps, cs, zs = read_basis(Z)
            ip = 0
            for L, k in zip(Ls, nps):
                c_sh = cs[ip:ip + k]
                z_sh = zs[ip:ip + k]
                ip += k
                for (l, m, n) in cartesian_shell(L):
                    istart.append(len(coef))
                    nprim.append(k)
                    cen.append(xyz[a])
                    powr.append((l, m, n))
                    bf_atom.append(a)
                    cu = [c_sh[t] * norm_prim(l, m, n, z_sh[t]) for t in range(k)]
                    if k > 1:   # fix_normalization (basis.f90:346-369)
                        nc = norm_contr(l, m, n, z_sh, cu)
                        cu = [x * nc for x in cu]
                    coef.extend(cu)
                    alpha.extend(z_sh)
        self.atno = np.array(atno, dtype=np.int32)
        self.xyz = xyz
        self.nbf = len(nprim)
        self.nprim = np.array(nprim, dtype=np.int32)
        self.istart = np.array(istart, dtype=np.int32)  # 0-based, as passed to C
        cen = np.array(cen, dtype=float).reshape(-1, 3)
        self.x = np.ascontiguousarray(cen[:, 0])
        self.y = np.ascontiguousarray(cen[:, 1])
        self.z = np.ascontiguousarray(cen[:, 2])
        powr = np.array(powr, dtype=np.int32).reshape(-1, 3)
        self.l = np.ascontiguousarray(powr[:, 0])
        self.m = np.ascontiguousarray(powr[:, 1])
        self.n = np.ascontiguousarray(powr[:, 2])
        self.coef = np.array(coef, dtype=float)
        self.alpha = np.array(alpha, dtype=float)
        self.bf_atom = np.array(bf_atom, dtype=np.int32)

    def cargs(self):
        return [self.nbf, _ptr(self.nprim), _ptr(self.istart), _ptr(self.x), _ptr(self.y), _ptr(self.z),
                _ptr(self.l), _ptr(self.m), _ptr(self.n), _ptr(self.coef), _ptr(self.alpha)]


def one_electron(B):
    """S, T, V (n x n, symmetric) as getS/getT/getV compute them (qc.c:6-93)."""
    n = B.nbf
    S = np.zeros(n * n)
    T = np.zeros(n * n)
    V = np.zeros(n * n)
    L = lib()
    L.orc_getS(*B.cargs(), _ptr(S))
    L.orc_getT(*B.cargs(), _ptr(T))
    xa = np.ascontiguousarray(B.xyz[:, 0])
    ya = np.ascontiguousarray(B.xyz[:, 1])
    za = np.ascontiguousarray(B.xyz[:, 2])
    L.orc_getV(*B.cargs(), len(B.atno), _ptr(B.atno), _ptr(xa), _ptr(ya), _ptr(za), _ptr(V))
    f = lambda a: a.reshape(n, n, order="F")
    return f(S), f(T), f(V)


def get_ints(B):
    """Packed 8-fold int2 (qc.c:95-132), 0-based ijkl2intindex."""
    n = B.nbf
    m = n * (n + 1) // 2
    out = np.zeros(m * (m + 1) // 2)
    lib().orc_getInts(*B.cargs(), _ptr(out))
    return out


def intindex(i, j, k, l):
    """0-based packed index (cints.c:327-347)."""
    if i < j:
        i, j = j, i
    if k < l:
        k, l = l, k
    ij = i * (i + 1) // 2 + j
    kl = k * (k + 1) // 2 + l
    if ij < kl:
        ij, kl = kl, ij
    return ij * (ij + 1) // 2 + kl


def libint_shell_walk(power):
    """Shells as the libint plug-in recovers them from the per-function powers
    (gaussians_libint.f90:70-124): from function i, lambda = sum(power(:, i)),
    the next shell starts (lambda+1)(lambda+2)/2 functions later.
    Returns [(first function, lambda)], 0-based."""
    power = np.asarray(power)
    out, i = [], 0
    while i < power.shape[1]:
        lam = int(power[:, i].sum())
        out.append((i, lam))
        i += (lam + 1) * (lam + 2) // 2
    return out


def unpack_int2(int2, n):
    """Full (i j|k l) array [n, n, n, n] from the packed 8-fold int2 (0-based
    index of cints.c:327-347)."""
    i = np.arange(n)
    I, J = np.meshgrid(i, i, indexing="ij")
    ij = np.where(I >= J, I * (I + 1) // 2 + J, J * (J + 1) // 2 + I)
    a = ij[:, :, None, None]
    b = ij[None, None, :, :]
    hi = np.maximum(a, b)
    lo = np.minimum(a, b)
    return int2[hi * (hi + 1) // 2 + lo]


def pack_int2(Y):
    """Packed 8-fold vector of a full [n, n, n, n] array with that symmetry."""
    n = Y.shape[0]
    m = n * (n + 1) // 2
    out = np.zeros(m * (m + 1) // 2)
    i = np.arange(n)
    I, J = np.meshgrid(i, i, indexing="ij")
    keep = I >= J
    ij = (I * (I + 1) // 2 + J)[keep]
    pairs = np.stack([I[keep], J[keep]], axis=1)
    a = ij[:, None]
    b = ij[None, :]
    sel = a >= b
    idx = (a * (a + 1) // 2 + b)[sel]
    pi = np.broadcast_to(np.arange(len(ij))[:, None], sel.shape)[sel]
    pk = np.broadcast_to(np.arange(len(ij))[None, :], sel.shape)[sel]
    out[idx] = Y[pairs[pi, 0], pairs[pi, 1], pairs[pk, 0], pairs[pk, 1]]
    return out


def transform_int2(int2, C):
    """mbpt.f90:11-79: orbital integrals (i j|k l) = sum C_mu,i C_nu,j C_al,k C_be,l
    (mu nu|al be), in the packed order of the input."""
    n = C.shape[0]
    T = unpack_int2(int2, n)
    Y = np.einsum("abcd,ai,bj,ck,dl->ijkl", T, C, C, C, C, optimize=True)
    return pack_int2(Y)


def mbpt2(moint2, lam, nocc):
    """Second-order energy, mbpt.f90:154-159 with mbpt2.f90:8-31:
    sum over occupied a, b and virtual r, s of
    [2 (ar|bs)(ra|sb) - (ar|bs)(sa|rb)] / (lam_a + lam_b - lam_r - lam_s)."""
    n = len(lam)
    Y = unpack_int2(moint2, n)
    o, v = slice(0, nocc), slice(nocc, n)
    vovo = Y[v, o, v, o]
    arbs = Y[o, v, o, v]                   # (a r|b s) at [a, r, b, s]
    rasb = vovo.transpose(1, 0, 3, 2)      # (r a|s b) at [a, r, b, s]
    sarb = vovo.transpose(1, 2, 3, 0)      # (s a|r b) at [a, r, b, s]
    lo, lv = lam[o], lam[v]
    den = lo[:, None, None, None] - lv[None, :, None, None] + lo[None, None, :, None] - lv[None, None, None, :]
    return float(np.sum((2.0 * arbs * rasb - arbs * sarb) / den))


def eri(B, i, j, k, l):
    return lib().orc_eri_bf(*B.cargs(), i, j, k, l)


def deri(B, i, j, k, l):
    """d(ij|kl)/dR for the centres of i, j, k, l: array [4 centres, 3 axes]."""
    out = np.zeros((3, 4))
    for ax in range(3):
        o = np.zeros(4)
        lib().orc_deri_bf(*B.cargs(), i, j, k, l, ax, _ptr(o))
        out[ax] = o
    return out.T


def eri_quartets_sum(B, first, nfun, nthreads=0):
    """Sum of all integrals of the given shell quartets (first/nfun: [nq, 4] arrays
    of first function and function count); used to time the THO restatement."""
    f = np.ascontiguousarray(first, dtype=np.int32)
    nf = np.ascontiguousarray(nfun, dtype=np.int32)
    return lib().orc_eri_quartets(*B.cargs(), len(f), _ptr(f), _ptr(nf), int(nthreads))


def jk(B, P, nthreads=0):
    """J_ij = sum P_kl (ij|kl), K_ij = sum P_kl (il|kj) (scf.f90:26-37)."""
    n = B.nbf
    Pf = np.asfortranarray(P, dtype=float)
    J = np.zeros((n, n), order="F")
    K = np.zeros((n, n), order="F")
    lib().orc_jk(*B.cargs(), _ptr(Pf), _ptr(J), _ptr(K), int(nthreads))
    return J, K


def shells(B):
    """(first function, function count) of every shell, in function order
    (the libint plug-in's walk over the per-function powers)."""
    walk = libint_shell_walk(np.stack([B.l, B.m, B.n]))
    first = np.array([f for f, _ in walk], dtype=np.int32)
    nf = np.array([(L + 1) * (L + 2) // 2 for _, L in walk], dtype=np.int32)
    return first, nf


def schwarz(B, nthreads=0):
    """Packed Schwarz bounds Q[A(A+1)/2 + B] = sqrt(max |(ab|ab)|) of all shell pairs A >= B."""
    first, nf = shells(B)
    ns = len(first)
    Q = np.zeros(ns * (ns + 1) // 2)
    lib().orc_schwarz(*B.cargs(), ns, _ptr(first), _ptr(nf), _ptr(Q))
    return Q


def jk_screened(B, P, thr=1e-12, nthreads=0, Q=None):
    """J, K over the unique shell quartets with Q_AB Q_CD >= thr (the direct
    engine's Schwarz screening), THO integrals; returns (J, K, quartets)."""
    n = B.nbf
    first, nf = shells(B)
    if Q is None:
        Q = schwarz(B, nthreads)
    Pf = np.asfortranarray(P, dtype=float)
    J = np.zeros((n, n), order="F")
    K = np.zeros((n, n), order="F")
    nq = lib().orc_jk_screened(*B.cargs(), len(first), _ptr(first), _ptr(nf), _ptr(np.ascontiguousarray(Q)),
                               float(thr), _ptr(Pf), _ptr(J), _ptr(K), int(nthreads))
    return J, K, nq


def formG(B, P, nthreads=0):
    J, K = jk(B, P, nthreads)
    return J - 0.5 * K


def nuclear_energy(atno, xyz):
    """scf.f90:138-150"""
    xyz = np.asarray(xyz, dtype=float).reshape(-1, 3)
    E = 0.0
    for i in range(len(atno)):
        for j in range(i):
            E += atno[i] * atno[j] / math.sqrt(float(np.dot(xyz[i] - xyz[j], xyz[i] - xyz[j])))
    return E


class SCFResult:
    pass


def doscf(H, G_of_P, S, nocc, Nscf, tolE, tolP, alpha, P0=None, log=None):
    """Restates doscf (scf.f90:72-136): P = 0 start, F = H + G(P),
    E0 = 1/2 sum P(H+F), eigh(F, S) by LAPACK dsygvd on the lower triangle,
    P = 2 C_occ C_occ^T, dP = sqrt(sum (P-Pold)^2 / 4), convergence when
    i > 3 (1-based) and dP < tolP and max-min of the last 4 E0 < tolE, else
    linear mixing P = (1-alpha) Pold + alpha P."""
    n = H.shape[0]
    P = np.zeros((n, n)) if P0 is None else np.array(P0, dtype=float)
    Escf = []
    for it in range(1, Nscf + 1):
        G = G_of_P(P)
        F = H + G
        E0 = 0.5 * float(np.sum(P * (H + F)))
        Escf.append(E0)
        if log is not None:
            log.append(E0)
        lam, C = scipy.linalg.eigh(F, S, lower=True, driver="gvd")
        Pold = P
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        dP = math.sqrt(float(np.sum((P - Pold) ** 2)) / 4)
        if it > 3:
            w = Escf[it - 4:it]
            if dP < tolP and max(w) - min(w) < tolE:
                r = SCFResult()
                r.C, r.P, r.lam, r.E0, r.F, r.G, r.iterations, r.dP = C, P, lam, E0, F, G, it, dP
                r.Pin = Pold
                r.E_iter = list(Escf)
                return r
        P = (1 - alpha) * Pold + alpha * P
    raise RuntimeError("SCF did not converge in %d iterations." % Nscf)


def rhf_gauss(atno, xyz, Nscf=40, alpha=1.0, tolE=1e-10, tolP=1e-8, nthreads=0, P0=None):
    """drivers.f90:12-51 (S, T, V and G(P) from the oracle integrals)."""
    B = Basis(atno, xyz)
    S, T, V = one_electron(B)
    H = T + V
    nelec = int(sum(atno))
    r = doscf(H, lambda P: formG(B, P, nthreads), S, nelec // 2, Nscf, tolE, tolP, alpha, P0=P0)
    r.Enuc = nuclear_energy(atno, xyz)
    r.Ekin = float(np.sum(r.P * T))
    r.Etot = r.E0 + r.Enuc
    r.S, r.T, r.V, r.H, r.basis = S, T, V, H, B
    return r


def energy_weighted_density(C, lam, nocc):
    """W = 2 sum_occ eps_i C_i C_i^T (the overlap-derivative weight of the RHF gradient)."""
    Co = C[:, :nocc]
    return 2.0 * (Co * lam[:nocc]) @ Co.T


def gradient(r, nthreads=0):
    """Analytic RHF nuclear gradient (3 x nat, Eh/bohr) at a converged oracle SCF result
    (no reference counterpart; checked against finite differences of rhf_gauss)."""
    B = r.basis
    nat = len(B.atno)
    nocc = int(sum(B.atno)) // 2
    P = np.asfortranarray(2.0 * r.C[:, :nocc] @ r.C[:, :nocc].T)
    W = np.asfortranarray(energy_weighted_density(r.C, r.lam, nocc))
    g2 = np.zeros(3 * nat)
    g1 = np.zeros(3 * nat)
    L = lib()
    L.orc_grad_2e(*B.cargs(), _ptr(B.bf_atom), nat, _ptr(P), _ptr(g2), int(nthreads))
    xa = np.ascontiguousarray(B.xyz[:, 0])
    ya = np.ascontiguousarray(B.xyz[:, 1])
    za = np.ascontiguousarray(B.xyz[:, 2])
    L.orc_grad_1e(*B.cargs(), _ptr(B.bf_atom), nat, _ptr(B.atno), _ptr(xa), _ptr(ya), _ptr(za), _ptr(P), _ptr(W),
                  _ptr(g1))
    gn = np.zeros((nat, 3))
    for i in range(nat):
        for j in range(nat):
            if i != j:
                d = B.xyz[i] - B.xyz[j]
                gn[i] -= B.atno[i] * B.atno[j] * d / np.linalg.norm(d) ** 3
    return (g1 + g2).reshape(nat, 3) + gn


# ------------------------------------------- the compiled reference itself ---
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")


def _ffloat(tok):
    """Fortran ES output drops the 'E' for 3-digit exponents (8.99-115)."""
    try:
        return float(tok)
    except ValueError:
        k = max(tok.rfind("-"), tok.rfind("+"))
        return float(tok[:k] + "E" + tok[k:])


def parse_ref_output(path):
    """Blocks "key count" + count values of ref_driver's output file; the int2
    sample (count -1) becomes int2_idx (1-based) and int2_val."""
    d = {}
    lines = open(path).read().split("\n")
    i = 0
    while i < len(lines):
        t = lines[i].split()
        if not t:
            i += 1
            continue
        key, cnt = t[0], int(t[1])
        i += 1
        if cnt == -1:
            rows = [r.split() for r in lines[i:] if r.strip()]
            d["int2_idx"] = np.array([[int(r[0]), int(r[1]), int(r[2]), int(r[3])] for r in rows], dtype=np.int32)
            d["int2_val"] = np.array([_ffloat(r[4]) for r in rows])
            break
        d[key] = np.array([_ffloat(x) for x in lines[i:i + cnt]])
        i += cnt
    return d


def run_reference(atno, xyz, Nscf=40, alpha=1.0, tolE=1e-10, tolP=1e-8, ndump=0, timeout=None):
    """Runs the compiled reference (oracle/_ref/ref_driver: gaussints, doscf,
    energies, drivers.f90:29-41) on one core in a temporary directory holding
    the 6-31G** file it reads (basis.f90:210). Returns (parsed output, the
    per-iteration E0 printed by doscf, scf.f90:114)."""
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError("oracle/_ref/ref_driver is not built (make -f oracle/Makefile.ref)")
    with tempfile.TemporaryDirectory(prefix="hfx_ref_") as work:
        shutil.copy(BASIS_FILE, os.path.join(work, "p631ss.txt"))
        inp = os.path.join(work, "case.in")
        out = os.path.join(work, "case.out")
        with open(inp, "w") as f:
            f.write("%d\n" % len(atno))
            for z, r in zip(atno, np.asarray(xyz, dtype=float).reshape(-1, 3)):
                f.write("%d %.17e %.17e %.17e\n" % (z, r[0], r[1], r[2]))
            f.write("%d %.17e %.17e %.17e\n%d\n" % (Nscf, alpha, tolE, tolP, ndump))
        env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
        # doscf keeps its n^4 TT array on the stack (scf.f90:79)
        log = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec %s %s %s" % (REF_DRIVER, inp, out)],
                             cwd=work, env=env, capture_output=True, text=True, check=True, timeout=timeout).stdout
        d = parse_ref_output(out)
    e_iter = []
    for line in log.split("\n"):
        t = line.split()
        if len(t) == 2 and t[0].isdigit():
            e_iter.append(float(t[1]))
    return d, e_iter
