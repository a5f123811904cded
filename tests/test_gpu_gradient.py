"""Analytic RHF forces on the GPU (derivative-ERI kernel + one-electron
derivative kernel): against the oracle's analytic gradient at the same density,
against central finite differences of the GPU SCF energy (h = 1e-4 bohr,
1e-6 Eh/bohr), translational invariance, the torch.autograd.Function over the
Cython binding (derivative kernels in backward), the Cython binding against the
ctypes one, and the mixed-precision gradient against the FP64 one."""
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


def test_cython_binding_matches_ctypes():
    """The Cython module (hfsolver_amd.hfx over c_hfx.pxd) and the ctypes
    binding call the same C ABI: equal S, T, V, G, J, K and gradient terms at
    the same inputs (FP64 atomics in another order: agreement to rounding)."""
    from hfsolver_amd import hfx
    atno, xyz = water_cluster(2, seed=2)
    b = Basis(atno, xyz)
    say("Cython vs ctypes binding, water dimer")
    ec, et = hfx.Engine(b), Engine(b)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((b.nbf, 10))
    P = X @ X.T / 10
    W = P @ P / b.nbf
    for A, B in zip(ec.one_electron(), et.one_electron()):
        assert np.abs(A - B).max() < 1e-14
    Gc, Jc, Kc, sc = ec.fock_2e(P, want_jk=True)
    Gt, Jt, Kt, stt = et.fock_2e(P, want_jk=True)
    assert sc.quartets_fp64 == stt.quartets_fp64 > 0
    for A, B in ((Gc, Gt), (Jc, Jt), (Kc, Kt)):
        assert np.abs(A - B).max() < 1e-12
    assert np.abs(ec.grad_2e(P) - et.grad_2e(P)).max() < 1e-12
    assert np.abs(ec.grad_1e(P, W) - et.grad_1e(P, W)).max() < 1e-12
    assert ec.grad_2e_stats().flops_fp64 > 0
    ec.close()
    et.close()
    with pytest.raises(hfx.HfxError):
        ec.fock_2e(P)


def test_autograd_backward_runs_the_derivative_kernels():
    """rhf_energy (Cython binding): forward runs the SCF only; the derivative
    kernels run in backward, on the engine and orbitals kept from forward, and
    the gradient equals rhf_gradient of the ctypes route."""
    import torch
    from hfsolver_amd import grad as gmod
    atno, xyz = water_cluster(2, seed=2)
    kw = dict(Nscf=120, tolE=1e-12, tolP=1e-10, diis=True)
    x = torch.tensor(np.array(xyz), dtype=torch.float64, requires_grad=True)
    gmod.last_backward_stats = None
    E = rhf_energy(x, atno, **kw)
    assert gmod.last_backward_stats is None
    E.backward()
    st = gmod.last_backward_stats
    assert st is not None and st.quartets_fp64 > 0 and st.jk_seconds > 0
    r = rhf_gauss(atno, np.array(xyz), keep_engine=True, **kw)
    g = rhf_gradient(r)
    r.engine.close()
    say("  |dE| %.1e, max|dg| %.1e" % (abs(E.item() - r.Etot), np.abs(x.grad.numpy() - g).max()))
    assert abs(E.item() - r.Etot) < 1e-10
    assert np.abs(x.grad.numpy() - g).max() < 1e-9
    # an engine kept by the caller is reused across calls (moved, not closed)
    from hfsolver_amd import hfx
    eng = hfx.Engine(Basis(atno, xyz))
    x2 = torch.tensor(np.array(xyz) + 0.01, dtype=torch.float64, requires_grad=True)
    E2 = rhf_energy(x2, atno, engine=eng, **kw)
    E2.backward()
    assert abs(E2.item() - E.item()) < 1e-8          # rigid translation
    assert np.abs(x2.grad.numpy() - x.grad.numpy()).max() < 1e-7
    assert eng.n == 50 and eng.screening()[1] > 0    # still open
    eng.close()
