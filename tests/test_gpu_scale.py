"""128 waters (n = 3200), twice the benchmark's cluster. The K rows of the
largest class combinations no longer fit in LDS there: by default those run as
narrow row launches (sliced over their a-components), with HFX_ROW_NARROW=0 as
tiles (jk_kernel), the rest in row mode either way. Both builds must equal the
all-tile build (kernel mode 1) to the atomics' rounding, with the same quartet
count, and their two-electron gradients (row-mode derivative launches on the
narrow combinations, the fallback kernel on the tiles) must agree; the
mixed-precision build must stay within |dE| < 1e-6 Eh of FP64.
Density: the core-Hamiltonian guess."""
import sys
import time

import numpy as np
import pytest

from hfsolver_amd.basis import Basis
from hfsolver_amd.clusters import water_cluster
from hfsolver_amd.engine import Engine

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def core_density(H, S, nocc):
    import torch
    dev = torch.device("cuda:0")
    Ht = torch.from_numpy(H).to(dev)
    St = torch.from_numpy(S).to(dev)
    L = torch.linalg.cholesky(St)
    Li = torch.linalg.solve_triangular(L, torch.eye(len(S), dtype=torch.float64, device=dev), upper=False)
    _, V = torch.linalg.eigh(Li @ Ht @ Li.T)
    C = Li.T @ V[:, :nocc]
    return (2.0 * C @ C.T).cpu().numpy()


def test_128_waters_row_and_tile_launches_agree(capfd, monkeypatch):
    atno, xyz = water_cluster(128, seed=128)
    b = Basis(atno, xyz)
    assert b.nbf == 3200
    say("128 waters: engines")
    t = time.time()
    eng = Engine(b, verbose=True)
    err = capfd.readouterr().err
    assert " rows:" in err and "narrow=1" in err and " tiles:" not in err
    try:
        S, T, V = eng.one_electron()
        P = core_density(T + V, S, int(sum(atno)) // 2)
        G, J, K, st = eng.fock_2e(P, want_jk=True)
        t = time.time()
        gn = np.asarray(eng.grad_2e(P))
        tg_narrow = time.time() - t
    finally:
        eng.close()
    monkeypatch.setenv("HFX_ROW_NARROW", "0")
    ew = Engine(b, verbose=True)
    err = capfd.readouterr().err
    assert " rows:" in err and " tiles:" in err and "narrow=1" not in err
    monkeypatch.delenv("HFX_ROW_NARROW")
    try:
        Gw, Jw, Kw, stw = ew.fock_2e(P, want_jk=True)
        t = time.time()
        gw = np.asarray(ew.grad_2e(P))
        say("  two-electron gradient, rows and tiles: %.2f s" % (time.time() - t))
    finally:
        ew.close()
    dg = np.abs(gn - gw).max()
    say("  gradient: max|dg| %.2e of max|g| %.2e" % (dg, np.abs(gw).max()))
    # (printed after the last capture of the engines' verbose output)
    say("  narrow rows: quartets %.4g, J/K %.3f s, two-electron gradient %.2f s" % (st.quartets_fp64, st.jk_seconds,
                                                                                   tg_narrow))
    assert dg <= 1e-10 * np.abs(gw).max()
    say("  rows and tiles: quartets %.4g, J/K %.3f s" % (stw.quartets_fp64, stw.jk_seconds))
    et = Engine(b, kernel_mode=1)
    try:
        Gt, Jt, Kt, stt = et.fock_2e(P, want_jk=True)
    finally:
        et.close()
    say("  all tiles: quartets %.4g, J/K %.3f s" % (stt.quartets_fp64, stt.jk_seconds))
    assert stt.quartets_fp64 == st.quartets_fp64 == stw.quartets_fp64
    for name, (Jx, Kx) in (("narrow rows", (J, K)), ("rows and tiles", (Jw, Kw))):
        dJ, dK = np.abs(Jx - Jt).max(), np.abs(Kx - Kt).max()
        say("  %s vs all tiles: max|dJ| %.2e max|dK| %.2e" % (name, dJ, dK))
        assert dJ <= 1e-12 * np.abs(Jt).max() and dK <= 1e-12 * np.abs(Kt).max()
    assert np.abs(G - G.T).max() < 1e-12
    em = Engine(b, fp32_threshold=1e-5)
    try:
        _, nq64, nq32 = em.screening()
        Gm = em.fock_2e(P)[0]
    finally:
        em.close()
    dE = 0.5 * float(np.sum(P * (Gm - G)))
    say("  mixed: FP32 share %.3f, |dE| %.2e Eh, max|dF| %.2e" % (nq32 / (nq32 + nq64), abs(dE),
                                                                np.abs(Gm - G).max()))
    assert abs(dE) < 1e-6
