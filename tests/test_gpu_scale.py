"""128 waters (n = 3200), twice the benchmark's cluster. The K rows of the
largest class combinations no longer fit in LDS there, so those run as tiles
(jk_kernel) while the rest stay in row mode: the build of that mixed
configuration must equal the all-tile build (kernel mode 1) to the atomics'
rounding, with the same quartet count; the mixed-precision build must stay
within |dE| < 1e-6 Eh of FP64. Density: the core-Hamiltonian guess."""
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


def test_128_waters_row_and_tile_launches_agree(capfd):
    atno, xyz = water_cluster(128, seed=128)
    b = Basis(atno, xyz)
    assert b.nbf == 3200
    say("128 waters: engines")
    t = time.time()
    eng = Engine(b, verbose=True)
    err = capfd.readouterr().err
    assert " rows:" in err and " tiles:" in err   # both launch forms in the default engine
    say("  set-up %.1f s" % (time.time() - t))
    try:
        S, T, V = eng.one_electron()
        P = core_density(T + V, S, int(sum(atno)) // 2)
        G, J, K, st = eng.fock_2e(P, want_jk=True)
    finally:
        eng.close()
    say("  default engine: quartets %.4g, J/K %.3f s" % (st.quartets_fp64, st.jk_seconds))
    et = Engine(b, kernel_mode=1)
    try:
        Gt, Jt, Kt, stt = et.fock_2e(P, want_jk=True)
    finally:
        et.close()
    say("  all tiles: quartets %.4g, J/K %.3f s" % (stt.quartets_fp64, stt.jk_seconds))
    assert stt.quartets_fp64 == st.quartets_fp64
    dJ, dK = np.abs(J - Jt).max(), np.abs(K - Kt).max()
    say("  max|dJ| %.2e max|dK| %.2e" % (dJ, dK))
    assert dJ <= 1e-12 * np.abs(J).max() and dK <= 1e-12 * np.abs(K).max()
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
