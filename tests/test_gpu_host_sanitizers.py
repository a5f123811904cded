"""The engine's host runtime under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/csrc/host_san is hfsolver_amd/csrc/hfx_host.cpp (shell regrouping,
primitive-pair screening, Schwarz-sorted class lists, ket tables, row-mode
block lists, every entry point's host buffers) built with both sanitizers for
the host, linked with the normal gfx950 kernel objects (built by build(),
tests/kernel_check.py). It drives every entry point of include/hfx.h on a
real basis: mixed-precision and tile-mode engines, density screening, moved
nuclei, explicit quartets with and without allocation, store mode, and the
rejected inputs (NaN in P, a short int2 array, inconsistent basis arrays).
Any invalid host access, leak of the engine's own allocations or undefined
behaviour ends it with a report; its G and gradient must equal the normal
library's. At 128 waters the (10) kets run in column windows, with their
full table for the gradient, and (22|10) in narrow rows.
"""
import os
import subprocess

import numpy as np
import pytest

from tests import kernel_check as KC

pytestmark = pytest.mark.gpu


def write_basis(path, b, P):
    n = b.nbf
    with open(path, "wb") as f:
        np.array([n, len(b.coef), len(b.atno)], dtype=np.int32).tofile(f)
        for a, t in ((b.nprim, np.int32), (b.istart, np.int32), (b.x, np.float64), (b.y, np.float64),
                     (b.z, np.float64), (b.l, np.int32), (b.m, np.int32), (b.n, np.int32), (b.coef, np.float64),
                     (b.alpha, np.float64), (b.atno, np.int32), (b.xyz, np.float64)):
            np.ascontiguousarray(a, dtype=t).tofile(f)
        np.asfortranarray(P, dtype=np.float64).ravel(order="F").tofile(f)


@pytest.mark.parametrize("nw", [2, 4, 128])
def test_host_runtime_under_asan_and_ubsan(nw, tmp_path):
    from hfsolver_amd.basis import Basis
    from hfsolver_amd.clusters import water_cluster
    from hfsolver_amd.engine import Engine

    if not os.path.exists(KC.HOST_SAN_EXE):
        pytest.fail("tests/csrc/host_san is not built (build() builds it)")
    atno, xyz = water_cluster(nw, seed=nw)
    b = Basis(atno, xyz)
    rng = np.random.default_rng(nw)
    X = rng.standard_normal((b.nbf, 5 * nw))
    P = X @ X.T / (5 * nw)
    bf, of = str(tmp_path / "basis.bin"), str(tmp_path / "out.bin")
    write_basis(bf, b, P)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:exitcode=23",
               LSAN_OPTIONS="suppressions=%s:print_suppressions=0" % KC.HOST_SAN_SUPP,
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    print("host_san: %d waters, n = %d" % (nw, b.nbf), flush=True)
    r = subprocess.run([KC.HOST_SAN_EXE, bf, of], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert "host_san: ok" in r.stdout
    # (the store mode, with its short-array check, runs up to n = 100)
    for what in ("NaN in P", "istart beyond") + (("short array",) if b.nbf <= 100 else ()):
        assert what in r.stdout
    out = np.fromfile(of, dtype=np.float64)
    n, nat = b.nbf, len(b.atno)
    G = out[:n * n].reshape(n, n, order="F")
    g = out[n * n:].reshape(nat, 3)
    eng = Engine(b, schwarz_threshold=1e-12, fp32_threshold=1e-5)
    try:
        G0 = eng.fock_2e(P)[0]
        g0 = eng.grad_2e(P)
    finally:
        eng.close()
    # the same engine code, summed by atomics in another order
    assert np.abs(G - G0).max() < 1e-12 * max(1.0, np.abs(G0).max())
    assert np.abs(g - g0).max() < 1e-11 * max(1.0, np.abs(g0).max())
