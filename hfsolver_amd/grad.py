"""Analytic RHF nuclear gradient and its torch.autograd.Function.

dE/dX = sum_ij P_ij dH_ij/dX - sum_ij W_ij dS_ij/dX                 (hfx_grad_1e, GPU)
      + sum_ijkl [1/2 P_ij P_kl - 1/4 P_ik P_jl] d(ij|kl)/dX          (hfx_grad_2e, GPU)
      + dE_nuc/dX,
with P = 2 C_occ C_occ^T and the energy-weighted density
W = 2 sum_occ eps_i C_i C_i^T of a converged SCF. hfsolver has no gradient
code; the result is checked against finite differences of the energy and
against the oracle's analytic gradient.
"""
import numpy as np

from .basis import Basis
from .engine import Engine
from .scf import nuclear_gradient, rhf_gauss


def energy_weighted_density(C, lam, nocc):
    Co = C[:, :nocc]
    return 2.0 * (Co * lam[:nocc]) @ Co.T


def rhf_gradient(result):
    """Gradient (nat x 3, Eh/bohr) of a converged rhf_gauss result that kept its engine."""
    eng = result.engine
    if eng is None:
        raise ValueError("rhf_gradient needs rhf_gauss(..., keep_engine=True)")
    atno = eng.basis.atno
    nocc = int(np.sum(atno)) // 2
    P = 2.0 * result.C[:, :nocc] @ result.C[:, :nocc].T
    W = energy_weighted_density(result.C, result.lam, nocc)
    return eng.grad_1e(P, W) + eng.grad_2e(P) + nuclear_gradient(atno, eng.basis.xyz)


class ForceField:
    """Repeated energy + force evaluations on one molecule, each SCF warm-started
    from the previous density (the engine is re-used through hfx_set_geometry)."""

    def __init__(self, atno, xyz, fp32_threshold=0.0, schwarz_threshold=1e-12, Nscf=60, tolE=1e-10, tolP=1e-8,
                 diis=True, alpha=1.0, eig="host", incremental=0.0, purify=False):
        self.atno = list(atno)
        self.engine = Engine(Basis(atno, xyz), schwarz_threshold=schwarz_threshold, fp32_threshold=fp32_threshold)
        self.kw = dict(Nscf=Nscf, tolE=tolE, tolP=tolP, diis=diis, alpha=alpha, eig=eig, incremental=incremental,
                       purify=purify)
        self.P = None
        self.last = None

    def _scf(self, xyz):
        xyz = np.asarray(xyz, dtype=float).reshape(-1, 3)
        if not np.array_equal(xyz, self.engine.basis.xyz):
            self.engine.set_geometry(xyz)
        r = rhf_gauss(self.atno, xyz, engine=self.engine, P0=self.P, **self.kw)
        self.P = r.P
        self.last = r
        return r

    def energy(self, xyz):
        """Total energy only (warm-started SCF)."""
        return self._scf(xyz).Etot

    def __call__(self, xyz):
        r = self._scf(xyz)
        return r.Etot, rhf_gradient(r)

    def close(self):
        self.engine.close()


def _torch_function():
    import torch

    class RHFEnergy(torch.autograd.Function):
        """E_tot(R) of closed-shell RHF; backward = analytic nuclear gradient."""

        @staticmethod
        def forward(ctx, xyz, atno, opts):
            x = xyz.detach().to("cpu", torch.float64).numpy()
            r = rhf_gauss(atno, x, keep_engine=True, **opts)
            try:
                g = rhf_gradient(r)
            finally:
                r.engine.close()
            ctx.save_for_backward(torch.from_numpy(g).to(dtype=xyz.dtype, device=xyz.device))
            return torch.tensor(r.Etot, dtype=xyz.dtype, device=xyz.device)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g, None, None

    return RHFEnergy


_RHF = None


def rhf_energy(xyz, atno, **opts):
    """Differentiable total energy: xyz is a float64 torch tensor (nat, 3) in bohr;
    opts are rhf_gauss keyword arguments (Nscf, alpha, tolE, tolP, diis, fp32_threshold, ...)."""
    global _RHF
    if _RHF is None:
        _RHF = _torch_function()
    return _RHF.apply(xyz, list(atno), opts)
