#!/usr/bin/env python
"""J+K Fock builds/s on a 64-water cluster in 6-31G** (n = 1600), one MI355X.

One step = one direct two-electron Fock build G(P) = J - K/2 (Schwarz
threshold 1e-12) at the converged FP64 SCF density, resident in HBM, through
the C ABI's hfx_fock_2e_device. The headline value is the mixed-precision form
(quartets with Q_ab Q_cd < 1e-5 in FP32 arithmetic, everything else FP64); the
all-FP64 form is timed beside it. Reported beside: |E_tot(mixed SCF) -
E_tot(FP64 SCF)|, the first-order |dE| = |1/2 tr P (G_mixed - G_fp64)| and
max |dF_ij| at the density, the rate with host buffers (PCIe included), the
analytic gradient (time and roofline of the derivative kernels), the
mixed-precision force workload (10 warm-started energy + force evaluations on
perturbed geometries, against the FP64 engine; the SCF densities by canonical
purification, one eigendecomposition per SCF), and the CPU baselines: the
oracle (THO restatement of the reference's integrals) on a sample of the same
quartets and on one full 16-water build, and the compiled reference's own
Fortran path on 1 and 2 waters (one core; it cannot run 16 or 64 waters).

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

PEAK_FP64 = 78.6e12   # MI355X vector FP64, FLOP/s
# The kernels are built without packed FP32 instructions: single-precision FMAs
# then issue at the FP64 rate, so 78.6 TFLOP/s is the peak for both parts of the
# mixed build (the packed-FP32 figure, 157.3, does not apply).
PEAK_VALU = PEAK_FP64


TRAFFIC_FILE = "profiles/jk_fock_64w.md"


def measured_traffic(key):
    """HBM bytes of one build of the 64-water workload (FETCH_SIZE + WRITE_SIZE
    of its J/K dispatches, scaled by a plain-copy calibration), as recorded by the
    rocprofv3 --pmc runs summarised in TRAFFIC_FILE under `key`
    (traffic_bytes_per_fp64_build / traffic_bytes_per_mixed_build); None when absent."""
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_FILE)):
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class JKTimedEngine:
    """An Engine whose device-resident Fock builds keep their J/K seconds
    (hfx_stats.jk_seconds, HIP events) in `jk`: the per-SCF-iteration J+K time."""

    def __init__(self, eng):
        self.eng, self.jk = eng, []

    def __getattr__(self, name):
        return getattr(self.eng, name)

    def fock_2e_device(self, *a, **kw):
        st = self.eng.fock_2e_device(*a, **kw)
        self.jk.append(st.jk_seconds)
        return st


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


def host_cpu():
    """CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def reference_cpu(nwaters):
    """The compiled reference's own RHF path (oracle/_ref/ref_driver: gaussints,
    then doscf with its n^4 TT tensor, src/drivers.f90:29-41, src/scf.f90:72-136)
    on one core for the given numbers of seeded waters. Above n = 120 (more than
    four waters) the reference stops (src/basis.f90:537): 16 and 64 waters are
    reported as not runnable."""
    from hfsolver_amd.clusters import water_cluster
    from oracle import oracle as O
    if not os.path.exists(O.REF_DRIVER):
        return {"status": "oracle/_ref/ref_driver not built"}
    runs = []
    for nw in nwaters:
        atno, xyz = water_cluster(nw, seed=nw)
        t0 = time.perf_counter()
        d, e_iter = O.run_reference(atno, xyz, Nscf=300, alpha=0.5, tolE=1e-10, tolP=1e-8)
        wall = time.perf_counter() - t0
        tg, ts, it = float(d["t_gaussints"][0]), float(d["t_doscf"][0]), len(e_iter)
        runs.append({"waters": nw, "nbf": int(d["n"][0]), "gaussints_s": tg, "doscf_s": ts, "scf_iterations": it,
                     "rhf_gauss_wall_s": wall, "fock_builds_per_s_stored_integrals": it / ts,
                     "first_fock_build_s_with_integrals": tg + ts / it})
        log("  reference Fortran, %d waters (n = %d): gaussints %.2f s, doscf %.2f s for %d iterations"
            % (nw, runs[-1]["nbf"], tg, ts, it))
    return {"runs": runs, "cores": 1, "threads": "MKL_NUM_THREADS=1 (the path has no OpenMP)",
            "16_waters": "n/a (cannot run: n = 400 > 120, src/basis.f90:537)",
            "64_waters": "n/a (cannot run: n = 1600 > 120, src/basis.f90:537)"}


def restatement_full_build(nwater=16, seed=16):
    """One complete, unextrapolated FP64 J/K build of nwater waters by the
    oracle (THO restatement of cints.c over the Schwarz-screened unique shell
    quartets, threshold 1e-12) on all host cores."""
    from hfsolver_amd.clusters import water_cluster
    from oracle import oracle as O
    atno, xyz = water_cluster(nwater, seed=seed)
    B = O.Basis(atno, xyz)
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B.nbf, 8))
    P = X @ X.T / 8
    threads = len(os.sched_getaffinity(0))
    Q = O.schwarz(B, threads)
    t0 = time.perf_counter()
    J, K, nq = O.jk_screened(B, P, thr=1e-12, nthreads=threads, Q=Q)
    t = time.perf_counter() - t0
    log("  CPU restatement, %d waters: one J/K build %.1f s on %d threads (%.0f quartets)" % (nwater, t, threads, nq))
    return {"waters": nwater, "nbf": B.nbf, "seconds": t, "builds_per_s": 1.0 / t, "threads": threads,
            "quartets": nq}, (atno, xyz, P)


def profile_builds(args):
    """One FP64 and one mixed-precision Fock build of the workload, nothing else
    timed: the run the rocprofv3 counter passes of the committed profiles use."""
    import scipy.linalg
    import torch

    from hfsolver_amd.basis import Basis
    from hfsolver_amd.clusters import water_cluster
    from hfsolver_amd.engine import Engine
    atno, xyz = water_cluster(args.nwater, seed=args.seed)
    b = Basis(atno, xyz)
    e = Engine(b)
    S, T, V = e.one_electron()
    nocc = int(sum(atno)) // 2
    lam, C = scipy.linalg.eigh(T + V, S, lower=True, driver="gvd")
    dP = torch.from_numpy(np.ascontiguousarray(2.0 * C[:, :nocc] @ C[:, :nocc].T)).to("cuda:0")
    dG = torch.zeros_like(dP)
    for name, thr in (("fp64", 0.0), ("mixed", args.fp32_threshold)):
        eng = Engine(b, fp32_threshold=thr, verbose=2 if args.verbose else 0)   # 2: per-class model FLOPs
        torch.cuda.synchronize()
        st = eng.fock_2e_device(dP, dG)
        log("  profile build %s: %.4f s (J/K kernels %.4f s)" % (name, st.seconds, st.jk_seconds))
        eng.close()
    e.close()
    return 0


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
    ap.add_argument("--ref-waters", default="1,2",
                    help="water counts of the reference Fortran timing (comma list, at most 4; '' skips it)")
    ap.add_argument("--no-full-cpu-build", action="store_true",
                    help="skip the complete 16-water build of the CPU restatement")
    ap.add_argument("--no-gradient", action="store_true", help="skip the analytic-gradient timing")
    ap.add_argument("--forces", type=int, default=10, metavar="N",
                    help="also run N consecutive energy+force evaluations (SCF warm-started from the previous "
                         "density) on seeded +-0.01 A perturbations, mixed engine vs FP64 engine (10 in the "
                         "mixed-precision force workload); reported under 'forces'")
    ap.add_argument("--incremental", type=float, default=0.0, metavar="THR",
                    help="force workload: incremental Fock builds with density-weighted screening at THR (0: off)")
    ap.add_argument("--eigenvectors", action="store_true",
                    help="force workload: SCF densities from an eigendecomposition every iteration instead of "
                         "canonical purification (matrix products; one eigendecomposition at convergence)")
    ap.add_argument("--verbose", action="store_true", help="per-class screening statistics on stderr")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (known bytes) to scale HBM counters")
    ap.add_argument("--profile-builds", action="store_true",
                    help="only run one FP64 and one mixed build (device buffers, core-Hamiltonian density) for "
                         "profiler counters; each build starts with one prep_kernel dispatch (with --verbose: "
                         "the modelled FLOPs of every class launch)")
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
    if args.profile_builds:
        return profile_builds(args)
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
            te = JKTimedEngine(eng)
            r = doscf(H, S, nocc, lambda X, e=eng: e.fock_2e(X)[0], eig="device", engine=te, **kw)
            scf[name] = {"Etot": r["E0"] + Enuc, "iterations": r["iterations"], "seconds": time.time() - t,
                         "jk_seconds_per_iteration_median": float(np.median(te.jk)),
                         "jk_seconds_per_iteration_max": float(np.max(te.jk)), "fock_builds": len(te.jk)}
            log("  %s SCF: E_tot %.10f Eh in %d iterations (%.1f s); J+K per iteration median %.4f s, max %.4f s"
                % (name, scf[name]["Etot"], r["iterations"], time.time() - t, np.median(te.jk), np.max(te.jk)))
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
    # algorithm-neutral lower bound of a build: sum_q K_q 3 nroots(L_q) N_q (the Rys assembly alone)
    lower64 = sum(e64.work_model()[1])
    lowermx = sum(emx.work_model()[1])
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
        gs64 = e64.grad_2e_stats()
        emx.grad_2e(Pc)
        t0 = time.perf_counter()
        emx.grad_2e(Pc)
        t2mx = time.perf_counter() - t0
        gsmx = emx.grad_2e_stats()
        t0 = time.perf_counter()
        e64.grad_1e(Pc, W)
        t1 = time.perf_counter() - t0
        log("  gradient: two-electron %.3f s (FP64), %.3f s (mixed); one-electron %.3f s" % (t2, t2mx, t1))
        g64 = gs64.flops_fp64 / gs64.jk_seconds
        gmx = (gsmx.flops_fp64 + gsmx.flops_fp32) / gsmx.jk_seconds
        grad = {"grad_2e_seconds_fp64": t2, "grad_2e_seconds_mixed": t2mx, "grad_1e_seconds": t1,
                "gradients_per_s_mixed": 1.0 / (t2mx + t1), "gradients_per_s_fp64": 1.0 / (t2 + t1),
                "roofline": {"kernel": "grad_row_kernel (all derivative launches of one gradient)",
                             "bound": "valu", "unit": "TFLOP/s", "peak": PEAK_VALU / 1e12,
                             "flops_fp64_gradient": gs64.flops_fp64, "kernel_seconds_fp64": gs64.jk_seconds,
                             "achieved_fp64": g64 / 1e12, "frac_fp64": g64 / PEAK_VALU,
                             "flops_mixed_gradient": gsmx.flops_fp64 + gsmx.flops_fp32,
                             "kernel_seconds_mixed": gsmx.jk_seconds, "achieved_mixed": gmx / 1e12,
                             "frac_mixed": gmx / PEAK_VALU}}
        log("  gradient kernels: FP64 %.3f s, %.1f TFLOP/s (%.3f of peak); mixed %.3f s, %.1f TFLOP/s"
            % (gs64.jk_seconds, g64 / 1e12, g64 / PEAK_VALU, gsmx.jk_seconds, gmx / 1e12))

    forces = None
    if args.forces > 0:
        from hfsolver_amd.clusters import ANG2BOHR
        from hfsolver_amd.grad import ForceField
        rng = np.random.Generator(np.random.PCG64(args.seed + 1))
        kw = dict(Nscf=80, tolE=1e-10, tolP=1e-7, diis=True, eig="device", incremental=args.incremental,
                  purify=not args.eigenvectors)
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
            Ef64, gf64 = ff64(x)
            d = np.asarray(gm) - np.asarray(gf64)
            gmax.append(float(np.abs(d).max()))
            grms.append(float(np.sqrt(np.mean(d * d))))
            dEs.append(abs(Em - Ef64))
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
                  "incremental_threshold": args.incremental, "purification": not args.eigenvectors}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(e64, b, st64[0].quartets_fp64)
        model, ncores = host_cpu()
        cpu["host"] = {"cpu_model": model, "cores_available": ncores}
        wl = [int(x) for x in args.ref_waters.split(",") if x.strip()]
        if any(w > 4 for w in wl):
            raise SystemExit("--ref-waters: the reference runs at most 4 waters (n <= 120)")
        if wl:
            cpu["reference_fortran"] = reference_cpu(wl)
        if not args.no_full_cpu_build:
            full, (atno16, xyz16, P16) = restatement_full_build()
            # the GPU FP64 engine on the same 16-water build
            e16 = Engine(Basis(atno16, xyz16))
            e16.fock_2e(P16)
            t0 = time.perf_counter()
            for _ in range(5):
                e16.fock_2e(P16)
            full["gpu_fp64_builds_per_s_host_buffers"] = 5.0 / (time.perf_counter() - t0)
            e16.close()
            cpu["restatement_full_build"] = full

    jk64 = np.mean([s.jk_seconds for s in st64])
    jkmx = np.mean([s.jk_seconds for s in stmx])
    f64 = st64[0].flops_fp64
    fmx64, fmx32 = stmx[0].flops_fp64, stmx[0].flops_fp32
    roof64 = f64 / jk64
    roofmx = (fmx64 + fmx32) / jkmx
    value = args.steps / tmx
    traffic = traffic64 = None
    if args.nwater == 64 and args.seed == 64 and abs(args.fp32_threshold - 1e-5) < 1e-20:
        traffic = measured_traffic("traffic_bytes_per_mixed_build")
        traffic64 = measured_traffic("traffic_bytes_per_fp64_build")
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
        "roofline": {"bound": "valu", "achieved": roofmx / 1e12, "peak": PEAK_VALU / 1e12,
                     "unit": "TFLOP/s", "frac": roofmx / PEAK_VALU, "traffic": traffic,
                     "kernel": "jk_row_kernel (all class launches of one mixed-precision build)",
                     "jk_seconds_mixed": float(jkmx), "flops_mixed": float(fmx64 + fmx32),
                     "flops_mixed_fp64_part": float(fmx64), "flops_mixed_fp32_part": float(fmx32),
                     "lower_bound_flops": lowermx, "lower_bound_frac": lowermx / jkmx / PEAK_VALU,
                     "fp64_build": {"achieved": roof64 / 1e12, "frac": roof64 / PEAK_FP64,
                                    "jk_seconds": float(jk64), "flops": float(f64), "traffic": traffic64,
                                    "lower_bound_flops": lower64, "lower_bound_frac": lower64 / jk64 / PEAK_FP64},
                     "traffic_source": TRAFFIC_FILE if (traffic or traffic64) else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
