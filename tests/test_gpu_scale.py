"""128 waters (n = 3200), twice the benchmark's cluster. The K rows of the
widest class combinations no longer fit in LDS there. By default the ket
tables of the widest pair classes are cut into column windows (columns in
spatial order, one row launch per window); with HFX_ROW_WINDOW_COLS=0 the
combinations whose rows do not fit run as narrow row launches (sliced over
their a-components), and with HFX_ROW_NARROW=0 as well as tiles (jk_kernel).
All three builds must equal the all-tile build (kernel mode 1) to the
atomics' rounding, with the same quartet count; their two-electron gradients
must agree; the mixed-precision build must stay within |dE| < 1e-6 Eh of
FP64. Density: the core-Hamiltonian guess."""
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


def test_128_waters_windows_rows_and_tiles_agree(capfd, monkeypatch):
    atno, xyz = water_cluster(128, seed=128)
    b = Basis(atno, xyz)
    assert b.nbf == 3200
    say("128 waters: engines")
    variants = [("column windows", {}, lambda e: "window 1/2" in e and " tiles:" not in e),
                ("narrow rows", {"HFX_ROW_WINDOW_COLS": "0"}, lambda e: "narrow=1" in e and " tiles:" not in e),
                ("rows and tiles", {"HFX_ROW_WINDOW_COLS": "0", "HFX_ROW_NARROW": "0"},
                 lambda e: " rows:" in e and " tiles:" in e and "narrow=1" not in e)]
    out, lines, P = {}, [], None
    for name, env, check in variants:
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        eng = Engine(b, verbose=True)
        for k in env:
            monkeypatch.delenv(k)
        err = capfd.readouterr().err
        assert check(err), name
        try:
            if P is None:
                S, T, V = eng.one_electron()
                P = core_density(T + V, S, int(sum(atno)) // 2)
            G, J, K, st = eng.fock_2e(P, want_jk=True)
            t = time.time()
            g = np.asarray(eng.grad_2e(P))
            tg = time.time() - t
        finally:
            eng.close()
        out[name] = (G, J, K, g, st)
        lines.append("  %s: quartets %.4g, J/K %.3f s, two-electron gradient %.2f s" % (name, st.quartets_fp64,
                                                                                       st.jk_seconds, tg))
    et = Engine(b, kernel_mode=1)
    try:
        Gt, Jt, Kt, stt = et.fock_2e(P, want_jk=True)
    finally:
        et.close()
    for line in lines:   # after the last capture of the engines' verbose output
        say(line)
    say("  all tiles: quartets %.4g, J/K %.3f s" % (stt.quartets_fp64, stt.jk_seconds))
    gref = out["rows and tiles"][3]
    for name, (G, J, K, g, st) in out.items():
        assert st.quartets_fp64 == stt.quartets_fp64
        dJ, dK = np.abs(J - Jt).max(), np.abs(K - Kt).max()
        dg = np.abs(g - gref).max()
        say("  %s vs all tiles: max|dJ| %.2e max|dK| %.2e; gradient max|dg| %.2e" % (name, dJ, dK, dg))
        assert dJ <= 1e-12 * np.abs(Jt).max() and dK <= 1e-12 * np.abs(Kt).max()
        assert dg <= 1e-10 * np.abs(gref).max()
        assert np.abs(G - G.T).max() < 1e-12
    G = out["column windows"][0]
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
