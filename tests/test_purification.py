"""Canonical purification (hfsolver_amd.scf.canonical_purification) on the CPU:
the projector onto the lowest eigenvectors of a gapped symmetric matrix,
against the eigendecomposition (the route of dsygvd in src/linalg.f90:241-280)."""
import numpy as np
import torch

from hfsolver_amd.scf import canonical_purification


def test_purification_matches_eigenprojector():
    rng = np.random.default_rng(3)
    n, nocc = 120, 37
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.concatenate([np.sort(rng.uniform(-20.0, -0.3, nocc)), np.sort(rng.uniform(0.2, 8.0, n - nocc))])
    A = (Q * ev) @ Q.T
    X = canonical_purification(torch.tensor(A), nocc).numpy()
    ref = Q[:, :nocc] @ Q[:, :nocc].T
    assert np.abs(X - ref).max() < 1e-10
    assert abs(np.trace(X) - nocc) < 1e-10
    assert np.abs(X @ X - X).max() < 1e-10
