"""Closed-shell RHF driver over the GPU two-electron engine.

rhf_gauss mirrors hfsolver's driver of the same name (src/drivers.f90:12-51)
and doscf (src/scf.f90:72-136): same arguments (atomic numbers, bohr
coordinates, Nscf, mixing alpha, tolE, tolP) and the same semantics:
  * P starts at 0, so the first Fock matrix is the core Hamiltonian (scf.f90:107);
  * each iteration: G = G(P) [here: one hfx_fock_2e call instead of
    formG(P, TT)], F = H + G, E0 = 1/2 sum P(H+F), F C = S C lambda by LAPACK
    dsygvd on the lower triangle (linalg.f90:241-280), P = 2 C_occ C_occ^T;
  * dP = sqrt(sum (P - Pold)^2 / 4); converged when the iteration count
    exceeds 3, dP < tolP and the spread of the last four E0 is below tolE;
    otherwise P = (1 - alpha) Pold + alpha P; failure to converge in Nscf
    iterations raises (the reference stops with exit code 1);
  * on return E0, lambda and C belong to the Fock matrix built from the last
    input density, P is the last output density; Enuc = sum Z_i Z_j / R_ij,
    Ekin = sum P T, Etot = E0 + Enuc.
Additions with no counterpart in the reference: optional DIIS extrapolation
of F (Pulay, error FPS - SPF), a warm-start density P0, the FP32 threshold of
the mixed-precision engine, and eig="device" (the iteration's linear algebra
on the GPU, the density kept in HBM).
"""
import math

import numpy as np
import scipy.linalg

from .basis import Basis
from .engine import Engine


class RHFResult:
    """Outputs of rhf_gauss (names follow the reference's arguments)."""


def nuclear_energy(atno, xyz):
    xyz = np.asarray(xyz, dtype=float).reshape(-1, 3)
    E = 0.0
    for i in range(len(atno)):
        for j in range(i):
            E += float(atno[i]) * float(atno[j]) / math.sqrt(float(np.sum((xyz[i] - xyz[j]) ** 2)))
    return E


def nuclear_gradient(atno, xyz):
    xyz = np.asarray(xyz, dtype=float).reshape(-1, 3)
    g = np.zeros_like(xyz)
    for i in range(len(atno)):
        for j in range(len(atno)):
            if i != j:
                d = xyz[i] - xyz[j]
                g[i] -= float(atno[i]) * float(atno[j]) * d / np.linalg.norm(d) ** 3
    return g


def _check_energy(it, E0):
    """A non-finite electronic energy (NaN or Inf in G or in the density) stops
    the SCF; the reference would carry on into LAPACK with it."""
    if not math.isfinite(E0):
        raise FloatingPointError("SCF iteration %d: non-finite electronic energy %r" % (it, E0))


def _diis_extrapolate(Fs, errs):
    m = len(Fs)
    B = -np.ones((m + 1, m + 1))
    B[m, m] = 0.0
    for i in range(m):
        for j in range(m):
            B[i, j] = float(np.vdot(errs[i], errs[j]))
    rhs = np.zeros(m + 1)
    rhs[m] = -1.0
    try:
        c = np.linalg.solve(B, rhs)[:m]
    except np.linalg.LinAlgError:
        return Fs[-1]
    return sum(ci * Fi for ci, Fi in zip(c, Fs))


def _doscf_device(H, S, nocc, engine, Nscf, alpha, tolE, tolP, diis, diis_size, P0, callback, device,
                  incremental=0.0, rebuild=0, purify=False):
    """doscf with the linear algebra on the GPU (torch on ROCm: rocBLAS / rocSOLVER
    in FP64) and the density kept in HBM (engine.fock_2e_device). F C = S C lam is
    solved as (L^-1 F L^-T) Y = Y lam, C = L^-T Y, with S = L L^T (Cholesky) -- the
    reduction dsygvd performs for the lower triangle (src/linalg.f90:241-280).

    incremental > 0: G(P) = G(P_prev) + G(P - P_prev), the difference built with
    density-weighted screening at that threshold (quartets with Q_ab Q_cd
    max|dP| below it are skipped); a full, unscreened build every `rebuild`
    iterations when rebuild > 0, and full builds once the density change is
    below 10 tolP, so that the last iterations (and the C, lam, P they return)
    carry no screening noise. At convergence E0, F and G are recomputed from one
    unscreened build of the final input density.

    purify: the density of each iteration by canonical purification of the
    orthogonalised Fock matrix (matrix products) instead of its eigenvectors;
    one eigendecomposition at convergence gives C and lam."""
    import torch
    dev = torch.device("cuda", device)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    n = H.shape[0]
    Ht, St = t(H), t(S)
    L = torch.linalg.cholesky(St)
    Li = torch.linalg.solve_triangular(L, torch.eye(n, dtype=torch.float64, device=dev), upper=False)
    P = torch.zeros((n, n), dtype=torch.float64, device=dev) if P0 is None else t(P0)
    G = torch.empty((n, n), dtype=torch.float64, device=dev)
    Escf, Fs, errs = [], [], []
    Pb, Gb = None, None   # density and G of the last build (incremental mode)
    dstate = {"B": np.zeros((0, 0)), "purify": purify}
    if incremental > 0:
        engine.set_density_screening(incremental)
    try:
        for it in range(1, Nscf + 1):
            # difference builds while the density still moves; full builds for the
            # last iterations, so that the energy test sees no screening noise
            near = dstate.get("dP", 1.0) < 10 * tolP
            if incremental > 0 and Pb is not None and not near and (rebuild <= 0 or (it - 1) % rebuild != 0):
                dG = torch.empty_like(G)
                engine.fock_2e_device((P - Pb).contiguous(), dG)
                G = Gb + dG
            else:
                # full builds carry no density screening: the last iterations and
                # the returned C, lam, P come from unscreened Fock matrices
                if incremental > 0:
                    engine.set_density_screening(0.0)
                G = torch.empty_like(G)
                engine.fock_2e_device(P, G)
                if incremental > 0:
                    engine.set_density_screening(incremental)
            Pb, Gb = P, G
            res = _device_step(it, P, G, Ht, St, Li, nocc, Escf, Fs, errs, diis, diis_size, callback, alpha, tolE, tolP,
                               dstate)
            if isinstance(res, dict):
                if incremental > 0:
                    engine.set_density_screening(0.0)
                    Pin = t(res["Pin"])
                    Gx = torch.empty_like(G)
                    engine.fock_2e_device(Pin, Gx)
                    F = Ht + Gx
                    res.update(E0=0.5 * float(torch.sum(Pin * (Ht + F))), F=F.cpu().numpy(), G=Gx.cpu().numpy())
                return res
            P = res
    finally:
        if incremental > 0:
            engine.set_density_screening(0.0)
    raise RuntimeError("SCF did not converge in %d iterations." % Nscf)


def canonical_purification(A, nocc, tol=1e-13, maxit=200):
    """Projector onto the nocc lowest eigenvectors of the symmetric A (orthogonal
    basis) by canonical purification (Palser and Manolopoulos, Phys. Rev. B 58,
    12704 (1998)): matrix products only, trace nocc throughout. The start
    X0 = (lam/n)(mu I - A) + (nocc/n) I maps the Gershgorin interval of A into
    [0, 1]; each step X <- ((1 - 2c) X + (1 + c) X^2 - X^3)/(1 - c) for c <= 1/2,
    ((1 + c) X^2 - X^3)/c otherwise (both keep [0, 1] in place and the trace),
    with c = tr(X^2 - X^3)/tr(X - X^2), until
    tr(X - X^2) (the idempotency error) is below tol n."""
    import torch
    n = A.shape[0]
    d = torch.diagonal(A)
    r = torch.sum(torch.abs(A), dim=1) - torch.abs(d)
    hmax, hmin = float(torch.max(d + r)), float(torch.min(d - r))
    mu = float(torch.sum(d)) / n
    lam = min(nocc / (hmax - mu), (n - nocc) / (mu - hmin))
    X = (-lam / n) * A
    X.diagonal().add_(lam * mu / n + nocc / n)
    for _ in range(maxit):
        X2 = X @ X
        X3 = X2 @ X
        t = torch.stack([torch.trace(X - X2), torch.trace(X2 - X3)]).cpu().numpy()
        if t[0] <= tol * n:
            break
        c = t[1] / t[0]
        if c <= 0.5:
            X = ((1 - 2 * c) * X + (1 + c) * X2 - X3) / (1 - c)
        else:
            X = ((1 + c) * X2 - X3) / c
    else:
        raise RuntimeError("canonical purification did not converge")
    return 0.5 * (X + X.T)


def _device_step(it, P, G, Ht, St, Li, nocc, Escf, Fs, errs, diis, diis_size, callback, alpha, tolE, tolP, dstate):
    """One iteration of _doscf_device after the Fock build: the converged result
    (dict) or the next density. dstate["B"] keeps the inner products of the
    DIIS error vectors, so that each iteration adds one row (one transfer)."""
    import torch
    F = Ht + G
    E0 = 0.5 * float(torch.sum(P * (Ht + F)))
    _check_energy(it, E0)
    Escf.append(E0)
    if callback is not None:
        callback(it, E0, P, G)
    Fd = F
    if diis and bool(torch.any(P != 0)):
        err = F @ P @ St - St @ P @ F
        Fs.append(F)
        errs.append(err)
        row = torch.stack([torch.sum(e * err) for e in errs]).cpu().numpy()
        Bo = dstate["B"]
        m = len(errs)
        Bn = np.zeros((m, m))
        Bn[:m - 1, :m - 1] = Bo
        Bn[m - 1, :] = row
        Bn[:, m - 1] = row
        if len(Fs) > diis_size:
            Fs.pop(0)
            errs.pop(0)
            Bn = Bn[1:, 1:]
        dstate["B"] = Bn
        if len(Fs) >= 2:
            m = len(Fs)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            B[:m, :m] = Bn
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            try:
                c = np.linalg.solve(B, rhs)[:m]
                Fd = sum(float(ci) * Fi for ci, Fi in zip(c, Fs))
            except np.linalg.LinAlgError:
                pass
    A = Li @ Fd @ Li.T
    A = 0.5 * (A + A.T)
    Pold = P
    if dstate.get("purify"):
        X = canonical_purification(A, nocc)
        P = 2.0 * Li.T @ X @ Li
        C = lam = None
    else:
        lam, Y = torch.linalg.eigh(A)
        C = Li.T @ Y
        Co = C[:, :nocc]
        P = 2.0 * Co @ Co.T
    dP = math.sqrt(float(torch.sum((P - Pold) ** 2)) / 4)
    dstate["dP"] = dP
    if it > 3:
        w = Escf[it - 4:it]
        if dP < tolP and max(w) - min(w) < tolE:
            if C is None:   # orbitals and orbital energies of the converged operator
                lam, Y = torch.linalg.eigh(A)
                C = Li.T @ Y
                Co = C[:, :nocc]
                P = 2.0 * Co @ Co.T
            cpu = lambda x: x.cpu().numpy()
            return dict(C=cpu(C), P=cpu(P), Pin=cpu(Pold), lam=cpu(lam), E0=E0, F=cpu(F), G=cpu(G),
                        iterations=it, dP=dP, E_iter=Escf)
    if not diis:
        P = (1 - alpha) * Pold + alpha * P
    return P


def doscf(H, S, nocc, fock_2e, Nscf=40, alpha=1.0, tolE=1e-10, tolP=1e-8, diis=False, diis_size=8, P0=None,
          callback=None, eig="host", engine=None, device=0, incremental=0.0, purify=False):
    """The SCF loop of src/scf.f90:72-136 with G(P) from `fock_2e`.

    eig="host": LAPACK dsygvd on the lower triangle as the reference does;
    eig="device": the same steps on the GPU (needs `engine`, whose
    fock_2e_device is used instead of fock_2e); there incremental > 0 builds
    G(P) = G(P_prev) + G(P - P_prev) with density-weighted screening at that
    threshold (no counterpart in the reference)."""
    if eig == "device":
        if engine is None:
            raise ValueError("doscf(eig='device') needs the engine")
        return _doscf_device(H, S, nocc, engine, Nscf, alpha, tolE, tolP, diis, diis_size, P0, callback, device,
                             incremental, purify=purify)
    if eig != "host":
        raise ValueError("eig must be 'host' or 'device'")
    if incremental or purify:
        raise ValueError("incremental builds and purification need eig='device'")
    n = H.shape[0]
    P = np.zeros((n, n)) if P0 is None else np.array(P0, dtype=float)
    Escf, Fs, errs = [], [], []
    for it in range(1, Nscf + 1):
        G = fock_2e(P)
        F = H + G
        E0 = 0.5 * float(np.sum(P * (H + F)))
        _check_energy(it, E0)
        if not np.all(np.isfinite(G)):
            raise FloatingPointError("SCF iteration %d: the two-electron matrix has non-finite elements" % it)
        Escf.append(E0)
        if callback is not None:
            callback(it, E0, P, G)
        Fd = F
        if diis and np.any(P):
            # the P = 0 start has a zero commutator and must not enter the subspace
            err = F @ P @ S - S @ P @ F
            Fs.append(F)
            errs.append(err)
            if len(Fs) > diis_size:
                Fs.pop(0)
                errs.pop(0)
            if len(Fs) >= 2:
                Fd = _diis_extrapolate(Fs, errs)
        lam, C = scipy.linalg.eigh(Fd, S, lower=True, driver="gvd")
        Pold = P
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        dP = math.sqrt(float(np.sum((P - Pold) ** 2)) / 4)
        if it > 3:
            w = Escf[it - 4:it]
            if dP < tolP and max(w) - min(w) < tolE:
                return dict(C=C, P=P, Pin=Pold, lam=lam, E0=E0, F=F, G=G, iterations=it, dP=dP, E_iter=Escf)
        if not diis:
            P = (1 - alpha) * Pold + alpha * P
    raise RuntimeError("SCF did not converge in %d iterations." % Nscf)


def rhf_gauss(atno, xyz, Nscf=40, alpha=1.0, tolE=1e-10, tolP=1e-8, *, diis=False, P0=None, fp32_threshold=0.0,
              schwarz_threshold=1e-12, engine=None, keep_engine=False, device=0, int2=False, eig="host",
              incremental=0.0, purify=False):
    """Closed-shell RHF of a neutral molecule (src/drivers.f90:12-51).

    With int2=True the result also carries `int2`, every unique (ij|kl) in the
    packed order the reference's rhf_gauss returns (store mode, all primitive
    pairs kept), for consumers of the stored integrals."""
    basis = Basis(atno, xyz) if engine is None else engine.basis
    eng = engine if engine is not None else Engine(basis, schwarz_threshold=schwarz_threshold,
                                                   fp32_threshold=fp32_threshold, device=device)
    try:
        S, T, V = eng.one_electron()
        H = T + V
        nelec = int(np.sum(atno))
        r = doscf(H, S, nelec // 2, lambda P: eng.fock_2e(P)[0], Nscf, alpha, tolE, tolP, diis=diis, P0=P0,
                  eig=eig, engine=eng, device=device, incremental=incremental, purify=purify)
    except Exception:
        if engine is None and not keep_engine:
            eng.close()
        raise
    out = RHFResult()
    for k, v in r.items():
        setattr(out, k, v)
    out.S, out.T, out.V, out.H = S, T, V, H
    out.Enuc = nuclear_energy(atno, xyz)
    out.Ekin = float(np.sum(out.P * T))
    out.Etot = out.E0 + out.Enuc
    out.basis = basis
    out.int2 = None
    if int2:
        store = Engine(basis, schwarz_threshold=0.0, device=device)
        try:
            out.int2 = store.int2()
        finally:
            store.close()
    if engine is None and not keep_engine:
        eng.close()
        out.engine = None
    else:
        out.engine = eng
    return out
