"""Cartesian Gaussian basis in hfsolver's per-basis-function layout.

Builds exactly the arrays gaussints hands to the integral code
(src/basis.f90:490-567): for every contracted Cartesian function its
primitive count, 0-based primitive offset, centre, powers (l, m, n) in
get_cartesian_shell order (src/basis.f90:371-409), normalised contraction
coefficients and exponents. Normalisation follows src/basis.f90:93-132 and
:425-488: each primitive gets norm_prim(l, m, n, alpha), and a contraction of
more than one primitive is then renormalised to unit self-overlap
(norm_contr). norm_prim's factor 2**(2L+1.5) is evaluated in single precision,
as the reference's default-real literal 1.5 makes it; the energy does not
depend on it, but matrix elements of uncontracted functions do at 1e-8.
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BASIS_FILE = os.path.join(HERE, "data", "p631ss.txt")
_L = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4, "h": 5, "i": 6}
_cache = {}


def _load(path=BASIS_FILE):
    if path in _cache:
        return _cache[path]
    data = {}
    with open(path) as f:
        tok = [t.split() for t in f if t.strip()]
    i = 0
    while i < len(tok):
        Z, nsh = int(tok[i][0]), int(tok[i + 1][0])
        i += 2
        shells = []
        for _ in range(nsh):
            L, k = _L[tok[i][0].lower()], int(tok[i][1])
            i += 1
            a = np.array([float(t[0]) for t in tok[i:i + k]])
            c = np.array([float(t[1]) for t in tok[i:i + k]])
            i += k
            shells.append((L, a, c))
        data[Z] = shells
    _cache[path] = data
    return data


def cartesian_powers(L):
    return [(L - k, k - n, n) for k in range(L + 1) for n in range(k + 1)]


def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


class Basis:
    """Per-function arrays of a molecule's 6-31G** basis (bohr)."""

    def __init__(self, atno, xyz, path=BASIS_FILE):
        data = _load(path)
        self.atno = np.ascontiguousarray(atno, dtype=np.int32)
        self.xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        nprim, istart, atom, powr, coef, alpha = [], [], [], [], [], []
        for ia, Z in enumerate(self.atno):
            if int(Z) not in data:
                raise ValueError("read_basis: atom Z=%d not found" % Z)
            for (L, a, c) in data[int(Z)]:
                # alpha-dependent part of norm_prim, with the FP32 factor 2**(2L+1.5)
                two = float(np.float32(2.0) ** np.float32(2 * L + 1.5))
                for (l, m, n) in cartesian_powers(L):
                    f = _dfact(2 * l - 1) * _dfact(2 * m - 1) * _dfact(2 * n - 1)
                    cu = c * np.sqrt(two * a ** (L + 1.5) / f / math.pi ** 1.5)
                    if len(a) > 1:
                        S = np.sum(np.outer(cu, cu) / np.add.outer(a, a) ** (L + 1.5))
                        cu = cu * (1 / math.pi ** 0.75 * math.sqrt(2 ** L / f) / math.sqrt(S))
                    istart.append(len(coef))
                    nprim.append(len(a))
                    atom.append(ia)
                    powr.append((l, m, n))
                    coef.extend(cu.tolist())
                    alpha.extend(a.tolist())
        self.nbf = len(nprim)
        self.nprim = np.array(nprim, dtype=np.int32)
        self.istart = np.array(istart, dtype=np.int32)
        self.bf_atom = np.array(atom, dtype=np.int32)
        p = np.array(powr, dtype=np.int32).reshape(-1, 3)
        self.l = np.ascontiguousarray(p[:, 0])
        self.m = np.ascontiguousarray(p[:, 1])
        self.n = np.ascontiguousarray(p[:, 2])
        self.coef = np.array(coef)
        self.alpha = np.array(alpha)
        self.move(self.xyz)

    def move(self, xyz):
        self.xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        c = self.xyz[self.bf_atom]
        self.x = np.ascontiguousarray(c[:, 0])
        self.y = np.ascontiguousarray(c[:, 1])
        self.z = np.ascontiguousarray(c[:, 2])

    @property
    def nelec(self):
        return int(self.atno.sum())
