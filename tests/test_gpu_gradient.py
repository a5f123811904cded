"""Analytic RHF forces on the GPU (derivative-ERI kernel + host one-electron
terms): against the oracle's analytic gradient at the same density, against
central finite differences of the GPU SCF energy (h = 1e-4 bohr, 1e-6 Eh/bohr),
translational invariance, the torch.autograd.Function, and the mixed-precision
gradient against the FP64 one."""
import sys

import numpy as np
import pytest

from hfsolver_amd.basis import Basis
from hfsolver_amd.clusters import water_cluster, water_monomer
from hfsolver_amd.engine import Engine
from hfsolver_amd.grad import ForceField, energy_weighted_density, rhf_energy, rhf_gradient
from hfsolver_amd.scf import rhf_gauss
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_gradient_terms_match_oracle_at_same_density():
    atno, xyz = water_monomer()
    say("gradient terms vs oracle, water")
    r = O.rhf_gauss(atno, xyz)
    nocc = 5
    P = np.asfortranarray(2.0 * r.C[:, :nocc] @ r.C[:, :nocc].T)
    W = np.asfortranarray(O.energy_weighted_density(r.C, r.lam, nocc))
    B = r.basis
    g2o = np.zeros(9)
    g1o = np.zeros(9)
    L = O.lib()
    L.orc_grad_2e(*B.cargs(), O._ptr(B.bf_atom), 3, O._ptr(P), O._ptr(g2o), 0)
    xa, ya, za = [np.ascontiguousarray(B.xyz[:, k]) for k in range(3)]
    L.orc_grad_1e(*B.cargs(), O._ptr(B.bf_atom), 3, O._ptr(B.atno), O._ptr(xa), O._ptr(ya), O._ptr(za),
                  O._ptr(P), O._ptr(W), O._ptr(g1o))
    eng = Engine(Basis(atno, xyz))
    g2 = eng.grad_2e(P)
    g1 = eng.grad_1e(P, W)
    eng.close()
    assert np.abs(g2.ravel() - g2o).max() < 1e-11
    assert np.abs(g1.ravel() - g1o).max() < 1e-11


def test_gradient_finite_differences_dimer():
    atno, xyz = water_cluster(2, seed=2)
    xyz = np.array(xyz)
    say("gradient vs finite differences, water dimer")
    kw = dict(Nscf=120, tolE=1e-12, tolP=1e-10, diis=True)
    r = rhf_gauss(atno, xyz, keep_engine=True, **kw)
    g = rhf_gradient(r)
    r.engine.close()
    assert np.abs(g.sum(axis=0)).max() < 1e-7
    rng = np.random.default_rng(3)
    comps = rng.choice(3 * len(atno), 12, replace=False)
    h = 1e-4
    for c in comps:
        a, d = divmod(int(c), 3)
        e = []
        for sgn in (+1, -1):
            x = xyz.copy()
            x[a, d] += sgn * h
            e.append(rhf_gauss(atno, x, **kw).Etot)
        fd = (e[0] - e[1]) / (2 * h)
        say("  component %d: analytic %.9f fd %.9f diff %.2e" % (c, g[a, d], fd, g[a, d] - fd))
        assert abs(fd - g[a, d]) < 1e-6


def test_autograd_function_matches_gradient():
    import torch
    atno, xyz = water_monomer()
    x = torch.tensor(np.array(xyz), dtype=torch.float64, requires_grad=True)
    E = rhf_energy(x, atno, Nscf=80, tolE=1e-12, tolP=1e-10)
    E.backward()
    r = rhf_gauss(atno, np.array(xyz), Nscf=80, tolE=1e-12, tolP=1e-10, keep_engine=True)
    g = rhf_gradient(r)
    r.engine.close()
    assert abs(E.item() - r.Etot) < 1e-10
    assert np.abs(x.grad.numpy() - g).max() < 1e-9


def test_mixed_precision_gradient_dimer():
    atno, xyz = water_cluster(2, seed=2)
    say("mixed-precision gradient, water dimer")
    f64 = ForceField(atno, xyz)
    f32 = ForceField(atno, xyz, fp32_threshold=1e-5)
    e64, g64 = f64(xyz)
    e32, g32 = f32(xyz)
    say("  |dE| %.2e  max|dg| %.2e" % (abs(e32 - e64), np.abs(g32 - g64).max()))
    assert abs(e32 - e64) < 1e-6
    assert np.abs(g32 - g64).max() < 1e-5
    # warm-started re-evaluation on a displaced geometry
    x2 = np.array(xyz) + 0.01
    e2, g2 = f64(x2)
    assert abs(e2 - e64) < 1e-6   # rigid translation leaves the energy unchanged
    assert np.abs(g2 - g64).max() < 1e-6
    f64.close()
    f32.close()


def test_gradient_fallback_kernel_matches_row_mode():
    """kernel_mode=1 runs the per-sublist tile / gradient kernels (the fallback
    where the row-mode LDS does not fit); same two-electron gradient."""
    atno, xyz = water_cluster(2, seed=2)
    say("two-electron gradient, row mode vs fallback kernels")
    r = rhf_gauss(atno, xyz, Nscf=80, tolE=1e-12, tolP=1e-10, diis=True, keep_engine=True)
    nocc = int(sum(atno)) // 2
    P = 2.0 * r.C[:, :nocc] @ r.C[:, :nocc].T
    g_rows = r.engine.grad_2e(P)
    r.engine.close()
    eng = Engine(Basis(atno, xyz), kernel_mode=1)
    g_tiles = eng.grad_2e(P)
    eng.close()
    assert np.abs(g_rows - g_tiles).max() < 1e-11
