"""The engine's Rys integral templates (hfsolver_amd/csrc/eri_rys.h), compiled
for the host, against the oracle's THO integrals (a restatement of
src/cints.c:72-149): every class combination of the 6-31G** s/p/d basis,
FP64 and FP32 arithmetic, plus the Rys tables against mpmath-free checks."""
import ctypes
import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O
from hfsolver_amd.clusters import water_cluster
from tests import kernel_check as KC

H2O = ([8, 1, 1], [[0.0, 0.0, 0.0], [1.809, 0.0, 0.0], [-0.453549, 1.751221, 0.0]])


@pytest.fixture(scope="module")
def lib():
    return KC.load()


def _oriented(A, B, C, D):
    if A["L"] < B["L"]:
        A, B = B, A
    if C["L"] < D["L"]:
        C, D = D, C
    if KC.pair_class(A["L"], B["L"]) < KC.pair_class(C["L"], D["L"]):
        A, B, C, D = C, D, A, B
    return A, B, C, D


def _ref_block(Bas, A, B, C, D):
    nc = lambda s: (s["L"] + 1) * (s["L"] + 2) // 2
    ref = np.zeros((nc(A), nc(B), nc(C), nc(D)))
    for a, b, c, d in itertools.product(*(range(nc(s)) for s in (A, B, C, D))):
        ref[a, b, c, d] = O.eri(Bas, A["f0"] + a, B["f0"] + b, C["f0"] + c, D["f0"] + d)
    return ref


def test_all_class_combinations_fp64(lib):
    Bas = O.Basis(*H2O)
    sh = KC.shells_of(Bas)
    seen = {}
    for (i, j, k, l) in itertools.product(range(len(sh)), repeat=4):
        A, B, C, D = sh[i], sh[j], sh[k], sh[l]
        if not (A["L"] >= B["L"] and C["L"] >= D["L"] and
                KC.pair_class(A["L"], B["L"]) >= KC.pair_class(C["L"], D["L"])):
            continue
        key = (A["L"], B["L"], C["L"], D["L"])
        if seen.get(key, 0) >= 2:
            continue
        seen[key] = seen.get(key, 0) + 1
        v = KC.quartet(lib, A, B, C, D)
        ref = _ref_block(Bas, A, B, C, D)
        assert np.abs(v - ref).max() < 5e-15 * max(1.0, np.abs(ref).max()), key
    assert len(seen) == 21


def test_random_quartets_dimer_fp64_fp32(lib):
    atno, xyz = water_cluster(2, seed=2)
    Bas = O.Basis(atno, xyz)
    sh = KC.shells_of(Bas)
    rng = np.random.default_rng(7)
    for _ in range(120):
        A, B, C, D = _oriented(*[sh[i] for i in rng.integers(0, len(sh), 4)])
        ref = _ref_block(Bas, A, B, C, D)
        v64 = KC.quartet(lib, A, B, C, D)
        v32 = KC.quartet(lib, A, B, C, D, fp32=True)
        scale = np.abs(ref).max()
        assert np.abs(v64 - ref).max() <= 1e-15 + 1e-14 * scale
        # FP32 arithmetic: relative error bounded by a few thousand ulps of float
        assert np.abs(v32 - ref).max() <= 1e-30 + 5e-6 * scale


def test_rys_rule_reproduces_boys_moments(lib):
    lib.chk_rys.restype = ctypes.c_double
    u = np.zeros(5)
    w = np.zeros(5)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    for nr in range(1, 6):
        for T in [0.0, 1e-8, 0.37, 1.0, 4.5, 17.3, 33.0, 35.99, 36.0, 45.99, 46.0, 59.99, 60.0, 85.0, 1e3, 1e5]:
            lib.chk_rys(nr, ctypes.c_double(T), p(u), p(w))
            F = O.boys(2 * nr - 1, T)
            for k in range(2 * nr):
                assert abs(np.sum(w[:nr] * u[:nr] ** k) - F[k]) <= 1e-15 * F[0] + 4e-15 * F[k]
            assert np.all(np.diff(u[:nr]) > 0) and np.all(u[:nr] > 0) and np.all(u[:nr] < 1)
            # single-precision Chebyshev form (FP32 kernels), at T rounded to float
            Tf = float(np.float32(T))
            Ff = O.boys(2 * nr - 1, Tf)
            lib.chk_rys(-20 - nr, ctypes.c_double(T), p(u), p(w))
            assert np.all(u[:nr] > 0) and np.all(u[:nr] < 1)
            for k in range(2 * nr):
                assert abs(np.sum(w[:nr] * u[:nr] ** k) - Ff[k]) <= 1e-6 * Ff[0] + 4e-6 * Ff[k], (nr, T, k)
            if nr <= 2:   # closed-form rules from the Boys Taylor table
                lib.chk_rys(-nr, ctypes.c_double(T), p(u), p(w))
                for k in range(2 * nr):
                    assert abs(np.sum(w[:nr] * u[:nr] ** k) - F[k]) <= 1e-15 * F[0] + 4e-15 * F[k]
                # single-precision form (FP32 kernels), at T rounded to float
                Tf = float(np.float32(T))
                Ff = O.boys(2 * nr - 1, Tf)
                lib.chk_rys(-10 - nr, ctypes.c_double(T), p(u), p(w))
                assert np.all(np.diff(u[:nr]) > 0) and np.all(u[:nr] > 0) and np.all(u[:nr] < 1)
                for k in range(2 * nr):
                    assert abs(np.sum(w[:nr] * u[:nr] ** k) - Ff[k]) <= 2e-7 * Ff[0] + 1e-6 * Ff[k], (nr, T, k)


def test_one_root_recursion_and_short_taylor_rules(lib):
    """The compact one-root table (F_top(T0), e^{-T0}; downward recursion to
    F_0) with the FP64 kernels' Taylor terms matches the oracle's Boys
    moments as tightly as the full table; with the FP32 kernels' five terms,
    and the five-term closed forms of one and two roots (the FP32 kernels'
    FP64 rules), within the truncation bound (1/64)^5/5! < 8e-12 (two roots:
    the root formulas amplify it about tenfold), at every table interval's
    worst offset and beyond the table."""
    lib.chk_rys.restype = ctypes.c_double
    u = np.zeros(5)
    w = np.zeros(5)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    Ts = [(k + 0.5) / 32 - 1e-12 for k in range(46 * 32)] + [0.0, 1e-9, 45.99, 46.0, 60.0, 1e3, 1e9]
    for code, nr, rel in ((-31, 1, 1e-15), (-41, 1, 2e-11), (-51, 1, 2e-11), (-52, 2, 2e-10)):
        for T in Ts:
            lib.chk_rys(code, ctypes.c_double(T), p(u), p(w))
            F = O.boys(2 * nr - 1, T)
            for k in range(2 * nr):
                assert abs(np.sum(w[:nr] * u[:nr] ** k) - F[k]) <= rel * F[0] + 8 * rel * F[k], (code, T, k)
            assert np.all(np.diff(u[:nr]) > 0) and np.all(u[:nr] > 0) and np.all(u[:nr] < 1)


def test_derivative_quartets_match_oracle(lib):
    """grad_rys.h instantiated on the host: sum gam_abcd d(ab|cd)/d(A, B, C) for
    every class combination against the oracle's derivative integrals
    (orc_deri_bf: the same THO restatement differentiated on each centre), in
    FP64 and FP32 arithmetic."""
    atno, xyz = water_cluster(2, seed=2)
    Bas = O.Basis(atno, xyz)
    sh = KC.shells_of(Bas)
    rng = np.random.default_rng(5)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    seen = set()
    for _ in range(40000):
        if len(seen) == 21:
            break
        A, B, C, D = _oriented(*[sh[i] for i in rng.integers(0, len(sh), 4)])
        key = (A["L"], B["L"], C["L"], D["L"])
        if key in seen:
            continue
        seen.add(key)
        nc = lambda s: (s["L"] + 1) * (s["L"] + 2) // 2
        dims = [nc(s) for s in (A, B, C, D)]
        gam = rng.standard_normal(dims)
        scale = np.einsum("a,b,c,d->abcd", A["sc"], B["sc"], C["sc"], D["sc"])
        gs = np.ascontiguousarray(gam * scale)
        ref = np.zeros((4, 3))
        for a, b, c, d in itertools.product(*(range(x) for x in dims)):
            ref += gam[a, b, c, d] * O.deri(Bas, A["f0"] + a, B["f0"] + b, C["f0"] + c, D["f0"] + d)
        m = np.abs(ref).max()
        for fp32, tol in ((0, 1e-12), (1, 2e-5)):
            g = np.zeros(9)
            args = [A["L"], B["L"], C["L"], D["L"]] + [p(s["cen"]) for s in (A, B, C, D)]
            for s in (A, B, C, D):
                args += [len(s["a"]), p(s["a"]), p(s["c"])]
            assert lib.chk_grad_quartet(*args, p(gs), fp32, p(g)) == 0
            assert np.abs(g.reshape(3, 3) - ref[:3]).max() <= 1e-14 + tol * m, (key, fp32)
    assert len(seen) == 21


def test_rys_rules_at_the_worst_table_offsets(lib):
    """Every interval of the tables at its worst point. One and two roots: the
    Boys Taylor table at T = (k + 1/2)/32 - 1e-12, the largest distance 1/64
    from a grid point, k over all 1472 points of [0, 46); three to five roots:
    the midpoints of all 480 polynomial intervals (width 1/8) of [0, 60). Moments
    against the oracle's Boys function; the zeroth within 8e-16 F_0, which a
    Taylor series one term short (about 1.6e-15 F_0 near T = 0) would miss."""
    lib.chk_rys.restype = ctypes.c_double
    u = np.zeros(5)
    w = np.zeros(5)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    cases = [(nr, -nr, [(k + 0.5) / 32 - 1e-12 for k in range(46 * 32)]) for nr in (1, 2)]
    cases += [(nr, nr, [(k + 0.5) / 8 for k in range(480)]) for nr in (3, 4, 5)]
    for nr, code, Ts in cases:
        for T in Ts:
            lib.chk_rys(code, ctypes.c_double(T), p(u), p(w))
            F = O.boys(2 * nr - 1, T)
            assert abs(np.sum(w[:nr]) - F[0]) <= 8e-16 * F[0], (nr, T)
            for k in range(1, 2 * nr):
                assert abs(np.sum(w[:nr] * u[:nr] ** k) - F[k]) <= 1e-15 * F[0] + 8e-15 * F[k], (nr, T, k)
