"""Edge geometries through the C ABI against the oracle's unscreened J and K
(THO restatement of cints.c over every quartet):
- one oxygen atom: every quartet on one centre, every class up to (dd|dd);
- one hydrogen atom: five functions, no d shells;
- the hydrogen molecule: s and p shells only, so the classes with a d shell
  have empty pair lists;
- two waters 60 bohr apart: the pairs joining them fall below the Schwarz
  threshold, so their lists hold the pairs of one molecule only while the
  Coulomb quartets between the molecules stay.
FP64 in both kernel modes, and the mixed form with the FP32 share on."""
import sys

import numpy as np
import pytest

from hfsolver_amd.basis import Basis
from hfsolver_amd.clusters import water_cluster
from hfsolver_amd.engine import Engine
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def geometry(name):
    if name == "O":
        return [8], [[0.0, 0.0, 0.0]]
    if name == "H":
        return [1], [[0.0, 0.0, 0.0]]
    if name == "H2":
        return [1, 1], [[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]]
    atno, xyz = water_cluster(2, seed=2)
    xyz = np.array(xyz, dtype=float)
    xyz[3:] += np.array([60.0, 0.0, 0.0])
    return atno, xyz.tolist()


@pytest.fixture(scope="module")
def oracle_jk():
    cache = {}

    def get(name):
        if name not in cache:
            atno, xyz = geometry(name)
            ob = O.Basis(atno, xyz)
            rng = np.random.default_rng(len(name))
            X = rng.standard_normal((ob.nbf, ob.nbf))
            P = (X + X.T) / 2
            J, K = O.jk(ob, P)
            cache[name] = (atno, xyz, P, J, K)
        return cache[name]
    return get


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name", ["O", "H", "H2", "far_waters"])
def test_edge_geometry_jk_matches_oracle(oracle_jk, name, mode):
    atno, xyz, P, Jr, Kr = oracle_jk(name)
    say("edge geometry", name, "kernel mode", mode, "n =", len(P))
    eng = Engine(Basis(atno, xyz), kernel_mode=mode)
    try:
        npairs, nq64, nq32 = eng.screening()
        G, J, K, _ = eng.fock_2e(P, want_jk=True)
    finally:
        eng.close()
    assert npairs > 0 and nq64 > 0 and nq32 == 0
    scale = max(1.0, np.abs(Jr).max(), np.abs(Kr).max())
    dJ, dK = np.abs(J - Jr).max(), np.abs(K - Kr).max()
    dG = np.abs(G - (Jr - 0.5 * Kr)).max()
    say("  pairs %d quartets %d; max|dJ| %.1e max|dK| %.1e max|dG| %.1e (scale %.1f)" % (npairs, nq64, dJ, dK, dG,
                                                                                          scale))
    assert dJ < 1e-12 * scale and dK < 1e-12 * scale and dG < 1e-12 * scale


@pytest.mark.parametrize("name", ["O", "H2", "far_waters"])
def test_edge_geometry_mixed_precision(oracle_jk, name):
    atno, xyz, P, Jr, Kr = oracle_jk(name)
    say("edge geometry, mixed precision", name)
    eng = Engine(Basis(atno, xyz), fp32_threshold=1e-5)
    try:
        _, nq64, nq32 = eng.screening()
        G = eng.fock_2e(P)[0]
    finally:
        eng.close()
    dG = np.abs(G - (Jr - 0.5 * Kr)).max()
    say("  FP32 quartets %d of %d; max|dG| %.1e" % (nq32, nq32 + nq64, dG))
    assert nq64 > 0
    assert dG < 1e-7
