#!/usr/bin/env python
"""J+K Fock builds/s on a 64-water cluster in 6-31G** (n = 1600), one MI355X.

One step = one direct two-electron Fock build G(P) = J - K/2 (Schwarz
threshold 1e-12) at the converged FP64 SCF density, resident in HBM, through
the C ABI's hfx_fock_2e_device. The headline value is the mixed-precision form
(quartets with Q_ab Q_cd < 1e-5 in FP32 arithmetic, everything else FP64); the
all-FP64 form is timed beside it. Reported beside: |E_tot(mixed SCF) -
E_tot(FP64 SCF)|, the first-order |dE| = |1/2 tr P (G_mixed - G_fp64)| and
max |dF_ij| at the density, the rate with host buffers (PCIe included), the
analytic-gradient times, and the oracle timed on a sample of the same
quartets (cpu_baseline).

    python bench.py [--steps K] [--warmup W] [--nwater 64] [--density scf|core]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X vector / matrix FP64, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X vector FP32 (packed), FLOP/s


TRAFFIC_FILE = "profiles/jk_fock_64w.md"


def measured_traffic():
    """HBM bytes of one FP64 build of the 64-water workload (FETCH_SIZE + WRITE_SIZE
    of its J/K dispatches, scaled by a plain-copy calibration), as recorded by the
    rocprofv3 --pmc runs summarised in TRAFFIC_FILE; None when absent."""
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_FILE)):
            if line.startswith("traffic_bytes_per_fp64_build:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(eng, basis, nq_total, target_s=12.0, seed=0):
    """The oracle (THO restatement of the reference's cints.c, all host threads)
    timed on a uniform random sample of this build's screened shell quartets;
    builds/s extrapolated by the quartet count."""
    from oracle import oracle as O
    A, B, Q = eng.pair_bounds()
    ns, f0, Ls, _ = eng.shells()
    qmax = Q.max()
    sig = np.nonzero(Q * qmax >= 1e-12)[0]
    rng = np.random.default_rng(seed)
    ob = O.Basis(basis.atno, basis.xyz)
    nc = lambda L: (L + 1) * (L + 2) // 2

    def sample(k):
        parts, have = [], 0
        while have < k:
            i = rng.choice(sig, 4 * k)
            j = rng.choice(sig, 4 * k)
            ok = Q[i] * Q[j] >= 1e-12
            parts.append(np.stack([A[i[ok]], B[i[ok]], A[j[ok]], B[j[ok]]], axis=1))
            have += int(ok.sum())
        q = np.concatenate(parts)[:k]
        ncv = np.array([nc(L) for L in range(3)], dtype=np.int32)
        return np.ascontiguousarray(f0[q], dtype=np.int32), np.ascontiguousarray(ncv[Ls[q]], dtype=np.int32)

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    k = 64
    while True:
        f, nf = sample(k)
        t0 = time.perf_counter()
        O.eri_quartets_sum(ob, f, nf, threads)
        t = time.perf_counter() - t0
        if t >= target_s or k >= 1 << 25:
            break
        k = int(k * min(16.0, max(2.0, 1.2 * target_s / max(t, 1e-3))))
    per_build = t * nq_total / k
    log("  cpu baseline: %d sampled quartets in %.2f s on %d threads -> %.1f s per build" % (k, t, threads, per_build))
    return {"value": 1.0 / per_build, "unit": "Fock builds/s", "cores": threads, "kind": "port",
            "sample": "%d uniformly sampled screened shell quartets of this build, every integral by the oracle's "
                      "THO restatement of cints.c (%.1f s), extrapolated to %.4g quartets" % (k, t, nq_total)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nwater", type=int, default=64)
    ap.add_argument("--seed", type=int, default=64)
    ap.add_argument("--density", choices=["core", "scf"], default="scf",
                    help="timing density: converged FP64 SCF density (default) or the core-Hamiltonian guess")
    ap.add_argument("--fp32-threshold", type=float, default=1e-5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-gradient", action="store_true", help="skip the analytic-gradient timing")
    ap.add_argument("--forces", type=int, default=0, metavar="N",
                    help="also run N consecutive energy+force evaluations (SCF warm-started from the previous "
                         "density) on seeded +-0.01 A perturbations, mixed engine vs FP64 engine (10 in the "
                         "mixed-precision force workload); reported under 'forces'")
    ap.add_argument("--incremental", type=float, default=0.0, metavar="THR",
                    help="force workload: incremental Fock builds with density-weighted screening at THR (0: off)")
    ap.add_argument("--purify", action="store_true",
                    help="force workload: SCF densities by canonical purification instead of eigenvectors")
    ap.add_argument("--verbose", action="store_true", help="per-class screening statistics on stderr")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (known bytes) to scale HBM counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = 1 << 30
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": 3}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine's path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured mode).", file=sys.stderr)
        return 2

    import scipy.linalg
    import torch

    from hfsolver_amd.basis import Basis
    from hfsolver_amd.clusters import water_cluster
    from hfsolver_amd.engine import Engine
    from hfsolver_amd.scf import doscf, nuclear_energy

    atno, xyz = water_cluster(args.nwater, seed=args.seed)
    b = Basis(atno, xyz)
    n = b.nbf
    log("bench: %d waters, n = %d" % (args.nwater, n))
    t = time.time()
    e64 = Engine(b, schwarz_threshold=1e-12, verbose=args.verbose)
    log("  FP64 engine set-up %.2f s; screening %s" % (time.time() - t, e64.screening()))
    t = time.time()
    emx = Engine(b, schwarz_threshold=1e-12, fp32_threshold=args.fp32_threshold)
    log("  mixed engine set-up %.2f s; screening %s" % (time.time() - t, emx.screening()))
    t = time.time()
    S, T, V = e64.one_electron()
    H = T + V
    log("  one-electron %.2f s" % (time.time() - t))
    nocc = int(sum(atno)) // 2
    Enuc = nuclear_energy(atno, xyz)
    scf = {}
    r64 = None
    if args.density == "scf":
        # FP64 and mixed-precision SCF from the same start (P = 0), DIIS, linear algebra on the GPU
        kw = dict(Nscf=80, tolE=1e-9, tolP=1e-6, diis=True)
        for name, eng in (("fp64", e64), ("mixed", emx)):
            t = time.time()
            r = doscf(H, S, nocc, lambda X, e=eng: e.fock_2e(X)[0], eig="device", engine=eng, **kw)
            scf[name] = {"Etot": r["E0"] + Enuc, "iterations": r["iterations"], "seconds": time.time() - t}
            log("  %s SCF: E_tot %.10f Eh in %d iterations (%.1f s)" % (name, scf[name]["Etot"], r["iterations"],
                                                                     time.time() - t))
            if name == "fp64":
                r64 = r
        P = r64["P"]
    else:
        lam, C = scipy.linalg.eigh(H, S, lower=True, driver="gvd")
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    dev = torch.device("cuda:0")
    dP = torch.from_numpy(np.ascontiguousarray(P)).to(dev)
    dG64 = torch.zeros(n * n, dtype=torch.float64, device=dev)
    dGmx = torch.zeros(n * n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def run(eng, dG, steps):
        # fock_2e_device returns after the stream has finished: per-step wall times
        sts, ts = [], []
        torch.cuda.synchronize()
        for _ in range(steps):
            t0 = time.perf_counter()
            sts.append(eng.fock_2e_device(dP, dG))
            ts.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        return float(np.sum(ts)), sts, ts

    for _ in range(args.warmup):
        run(e64, dG64, 1)
        run(emx, dGmx, 1)
    t64, st64, ts64 = run(e64, dG64, args.steps)
    log("  FP64: %.4f s / build (median %.4f)" % (t64 / args.steps, np.median(ts64)))
    tmx, stmx, tsmx = run(emx, dGmx, args.steps)
    log("  mixed: %.4f s / build (median %.4f)" % (tmx / args.steps, np.median(tsmx)))
    G64 = dG64.cpu().numpy().reshape(n, n)
    Gmx = dGmx.cpu().numpy().reshape(n, n)
    dE = 0.5 * float(np.sum(P * (Gmx - G64)))
    dF = float(np.abs(Gmx - G64).max())
    E64 = 0.5 * float(np.sum(P * (2 * H + G64)))
    # host buffers: P upload and G download included (PCIe)
    nh = min(args.steps, 3)
    emx.fock_2e(P)
    t0 = time.perf_counter()
    for _ in range(nh):
        emx.fock_2e(P)
    t_host = (time.perf_counter() - t0) / nh
    log("  mixed, host buffers: %.4f s / build" % t_host)

    grad = None
    if not args.no_gradient and r64 is not None:
        from hfsolver_amd.grad import energy_weighted_density
        W = energy_weighted_density(r64["C"], r64["lam"], nocc)
        Pc = 2.0 * r64["C"][:, :nocc] @ r64["C"][:, :nocc].T
        e64.grad_2e(Pc)
        t0 = time.perf_counter()
        e64.grad_2e(Pc)
        t2 = time.perf_counter() - t0
        emx.grad_2e(Pc)
        t0 = time.perf_counter()
        emx.grad_2e(Pc)
        t2mx = time.perf_counter() - t0
        t0 = time.perf_counter()
        e64.grad_1e(Pc, W)
        t1 = time.perf_counter() - t0
        log("  gradient: two-electron %.3f s (FP64), %.3f s (mixed); one-electron %.3f s" % (t2, t2mx, t1))
        grad = {"grad_2e_seconds_fp64": t2, "grad_2e_seconds_mixed": t2mx, "grad_1e_seconds": t1,
                "gradients_per_s_mixed": 1.0 / (t2mx + t1)}

    forces = None
    if args.forces > 0:
        from hfsolver_amd.clusters import ANG2BOHR
        from hfsolver_amd.grad import ForceField
        rng = np.random.Generator(np.random.PCG64(args.seed + 1))
        kw = dict(Nscf=80, tolE=1e-10, tolP=1e-7, diis=True, eig="device", incremental=args.incremental,
                  purify=args.purify)
        ffm = ForceField(atno, xyz, fp32_threshold=args.fp32_threshold, **kw)
        ff64 = ForceField(atno, xyz, fp32_threshold=0.0, **kw)
        if r64 is not None:   # the first evaluation starts from the base geometry's density
            ffm.P = ff64.P = np.asarray(r64["P"])
        tm, gmax, grms, dEs, its = [], [], [], [], []
        for k in range(args.forces):
            x = xyz + rng.uniform(-0.01, 0.01, size=xyz.shape) * ANG2BOHR
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            Em, gm = ffm(x)
            torch.cuda.synchronize()
            tm.append(time.perf_counter() - t0)
            E64, g64 = ff64(x)
            d = np.asarray(gm) - np.asarray(g64)
            gmax.append(float(np.abs(d).max()))
            grms.append(float(np.sqrt(np.mean(d * d))))
            dEs.append(abs(Em - E64))
            its.append(int(ffm.last.iterations))
            log("  forces %d: mixed SCF+gradient %.2f s (%d SCF iterations), |dE| %.2e Eh, max|dg| %.2e, rms %.2e Eh/bohr"
                % (k, tm[-1], its[-1], dEs[-1], gmax[-1], grms[-1]))
        ffm.close()
        ff64.close()
        forces = {"evaluations": args.forces, "components": int(3 * len(atno)),
                  "gradients_per_s_mixed": args.forces / sum(tm), "seconds_per_evaluation_mixed": sum(tm) / args.forces,
                  "scf_iterations_mixed": its, "force_error_max_vs_fp64": max(gmax),
                  "force_error_rms_vs_fp64": float(np.sqrt(np.mean(np.square(grms)))),
                  "abs_dE_max_vs_fp64": max(dEs),
                  "perturbation": "uniform +-0.01 A per coordinate about the base geometry, seed %d" % (args.seed + 1),
                  "incremental_threshold": args.incremental, "purification": bool(args.purify)}

    cpu = None if args.no_cpu_baseline else cpu_baseline(e64, b, st64[0].quartets_fp64)

    jk64 = np.mean([s.jk_seconds for s in st64])
    jkmx = np.mean([s.jk_seconds for s in stmx])
    f64 = st64[0].flops_fp64
    fmx64, fmx32 = stmx[0].flops_fp64, stmx[0].flops_fp32
    roof64 = f64 / jk64
    # mixed: fraction of the time-weighted peak (FP64 part at 78.6, FP32 part at 157.3 TFLOP/s)
    frac_mx = (fmx64 / PEAK_FP64 + fmx32 / PEAK_FP32) / jkmx
    value = args.steps / tmx
    traffic = None
    if args.nwater == 64 and args.seed == 64:
        traffic = measured_traffic()
    nint = stmx[0].integrals_fp64 + stmx[0].integrals_fp32
    out = {
        "metric": "J+K Fock builds/s, 64 H2O 6-31G**: mixed-FP32 vs FP64 GPU, |dE|<1e-6 Eh",
        "value": value,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic seeded water cluster (seed %d), %s density" %
                (args.seed, "converged FP64 SCF" if args.density == "scf" else "core-Hamiltonian"),
        "config": {"workload": "direct J+K Fock build, %d waters 6-31G** (n=%d), Schwarz 1e-12, FP32 below %g"
                               % (args.nwater, n, args.fp32_threshold),
                   "nbf": n, "quartets_fp64": stmx[0].quartets_fp64, "quartets_fp32": stmx[0].quartets_fp32},
        "median_ms_mixed": 1e3 * float(np.median(tsmx)),
        "fp64_builds_per_s": args.steps / t64,
        "median_ms_fp64": 1e3 * float(np.median(ts64)),
        "mixed_vs_fp64_speedup": t64 / tmx,
        "mixed_builds_per_s_host_buffers": 1.0 / t_host,
        "contracted_integrals_per_s_mixed": nint * args.steps / tmx,
        "scf_abs_dE_mixed_vs_fp64": abs(scf["mixed"]["Etot"] - scf["fp64"]["Etot"]) if scf else None,
        "scf": scf or None,
        "mixed_abs_dE_at_density": abs(dE),
        "mixed_max_dF": dF,
        "E_elec_fp64_at_density": E64,
        "gradient": grad,
        "forces": forces,
        "roofline": {"bound": "fp64-valu", "achieved": roof64 / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": roof64 / PEAK_FP64, "traffic": traffic,
                     "kernel": "jk_row_kernel/jk_kernel (all class launches of one FP64 build)",
                     "jk_seconds_fp64": float(jk64), "jk_seconds_mixed": float(jkmx),
                     "flops_fp64": float(f64), "traffic_source": TRAFFIC_FILE if traffic else None,
                     "mixed_frac_of_weighted_peak": frac_mx},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
