"""The screened quartet lists the GPU builds (hfsolver_amd/csrc/hfx_lists.hip)
against a direct count over the pair bounds: for every class combination
(bra class >= ket class), every bra pair of the class list (Q descending,
ties by pair index) takes the kets with Q_ab Q_cd >= threshold (same class:
up to and including itself), the FP32 share being those below the FP32 cut.
The engine's quartet counts must equal these counts exactly, for row-mode
and tile-mode engines and two threshold pairs."""
import numpy as np
import pytest

from hfsolver_amd.basis import Basis
from hfsolver_amd.clusters import water_cluster
from hfsolver_amd.engine import Engine

pytestmark = pytest.mark.gpu


def pair_class(la, lb):
    return 0 if la == 0 else ((1 if lb == 0 else 2) if la == 1 else 3 + lb)


def direct_counts(A, B, Q, L, thr, thr32):
    cls = np.array([pair_class(L[a], L[b]) for a, b in zip(A, B)])
    qmax = Q.max()
    sig = (Q * qmax >= thr) & (Q > 0)
    lists = {}
    for c in range(6):
        idx = np.nonzero(sig & (cls == c))[0]
        lists[c] = Q[idx[np.lexsort((idx, -Q[idx]))]]   # Q descending, ties by pair index
    n64 = n32 = 0
    for cb in range(6):
        for ck in range(cb + 1):
            qb, qk = lists[cb], lists[ck]
            if len(qb) == 0 or len(qk) == 0:
                continue
            asc = qk[::-1]
            for i, qi in enumerate(qb):
                km = len(asc) - np.searchsorted(asc, thr / qi, side="left")   # kets with Q >= thr / qi
                if cb == ck:
                    km = min(km, i + 1)
                k6 = len(asc) - np.searchsorted(asc, thr32 / qi, side="left") if thr32 > 0 else km
                k64 = min(k6, km)
                n64 += k64
                n32 += km - k64
    return int(sig.sum()), n64, n32


@pytest.mark.parametrize("thr,thr32", [(1e-12, 1e-5), (1e-10, 1e-3)])
@pytest.mark.parametrize("kernel_mode", [0, 1])
def test_quartet_lists_match_a_direct_count(thr, thr32, kernel_mode):
    atno, xyz = water_cluster(8, seed=8)
    b = Basis(atno, xyz)
    eng = Engine(b, schwarz_threshold=thr, fp32_threshold=thr32, kernel_mode=kernel_mode)
    try:
        A, B, Q = eng.pair_bounds()
        _, _, L, _ = eng.shells()
        npairs, nq64, nq32 = eng.screening()
    finally:
        eng.close()
    nsig, n64, n32 = direct_counts(A, B, Q, L, thr, thr32)
    assert (int(npairs), int(nq64), int(nq32)) == (nsig, n64, n32)
    assert n32 > 0 and n64 > 0


def test_work_model_matches_the_build_statistics():
    """hfx_work_model: its model FLOPs are those a Fock build reports, and the
    Rys-assembly lower bound (sum_q K_q 3 nroots N_q) lies below them; the
    Cython binding returns the same figures."""
    from hfsolver_amd import hfx
    atno, xyz = water_cluster(4, seed=4)
    b = Basis(atno, xyz)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((b.nbf, 10))
    P = X @ X.T / 10
    eng = Engine(b, schwarz_threshold=1e-12, fp32_threshold=1e-5)
    try:
        (m64, m32), (l64, l32) = eng.work_model()
        st = eng.fock_2e(P)[3]
    finally:
        eng.close()
    assert (m64, m32) == pytest.approx((st.flops_fp64, st.flops_fp32), rel=1e-12)
    assert 0 < l64 < m64 and 0 < l32 < m32
    ce = hfx.Engine(b, schwarz_threshold=1e-12, fp32_threshold=1e-5)
    try:
        (c64, c32), (cl64, cl32) = ce.work_model()
        assert (c64, c32, cl64, cl32) == (m64, m32, l64, l32)
    finally:
        ce.close()
