"""Write the golden fixtures from the compiled reference (run in the build
container only; /root/reference does not exist on the GPU box).

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

For each case it writes the input geometry, runs oracle/_ref/ref_driver (the
reference's own gaussints + doscf + get_nuclear_energy + kinetic_energy,
src/drivers.f90:29-41) in a temporary directory holding p631ss.txt
(src/basis.f90:210), and stores the outputs as
a compressed .npz (no pickles): energies, orbital energies, S, T, V, C, the
output density P, J(P), K(P), G(P) = J - K/2 from the reference's packed int2,
a strided sample of int2 with 1-based (i,j,k,l), and the per-iteration E0
printed by doscf (src/scf.f90:114).
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from hfsolver_amd.clusters import water_cluster, water_monomer  # noqa: E402
from oracle import oracle as O  # noqa: E402


def run_ref(atno, xyz, Nscf, alpha, tolE, tolP, ndump):
    """One run of the compiled reference: (parsed output, per-iteration E0)."""
    return O.run_reference(atno, xyz, Nscf, alpha, tolE, tolP, ndump)


def run_case(name, atno, xyz, Nscf=40, alpha=1.0, tolE=1e-10, tolP=1e-8, ndump=400):
    d, e_iter = run_ref(atno, xyz, Nscf, alpha, tolE, tolP, ndump)
    n = int(d["n"][0])
    res = dict(atno=np.array(atno, dtype=np.int32), xyz=np.array(xyz, dtype=float),
               params=np.array([Nscf, alpha, tolE, tolP]),
               E0=d["E0"][0], Enuc=d["Enuc"][0], Ekin=d["Ekin"][0], Etot=d["Etot"][0],
               lam=d["lam"], E_iter=np.array(e_iter),
               int2_idx=d.get("int2_idx", np.zeros((0, 4), np.int32)),
               int2_val=d.get("int2_val", np.zeros(0)))
    for k in ("S", "T", "V", "C", "P", "J", "K", "G"):
        res[k] = d[k].reshape(n, n, order="F")
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **res)
    print(name, "n =", n, "Etot =", res["Etot"], "iterations:", len(e_iter))


def run_fd(name, atno, xyz, ncomp=12, h=1e-4, seed=3, Nscf=400, alpha=0.5, tolE=1e-12, tolP=1e-10):
    """Central differences of the reference's total energy (E(x + h e_c) - E(x - h e_c)) / 2h
    on ncomp seeded components: the reference's own energy differences, against
    which the engine's analytic gradient is checked."""
    import numpy.random as npr
    xyz = np.array(xyz, dtype=float)
    comps = npr.default_rng(seed).choice(3 * len(atno), ncomp, replace=False)
    Ep, Em = [], []
    for c in comps:
        a, k = divmod(int(c), 3)
        for sgn, dst in ((+1, Ep), (-1, Em)):
            x = xyz.copy()
            x[a, k] += sgn * h
            d, _ = run_ref(atno, x, Nscf, alpha, tolE, tolP, 0)
            dst.append(d["Etot"][0])
        print(name, "component", int(c), "fd", (Ep[-1] - Em[-1]) / (2 * h), flush=True)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), atno=np.array(atno, dtype=np.int32), xyz=xyz,
                        comps=np.array(comps, dtype=np.int32), h=np.array(h), E_plus=np.array(Ep),
                        E_minus=np.array(Em), params=np.array([Nscf, alpha, tolE, tolP]))


def run_jk16(name="water16_jk", nthreads=0):
    """J and K of the 16-water cluster (seed 16, n = 400) at the core-Hamiltonian
    guess density, from the oracle's Schwarz-screened THO restatement (threshold
    1e-12). The reference itself cannot run this size (n <= 120, src/basis.f90:537);
    the oracle is pinned to it at <= 4 waters by tests/test_oracle.py."""
    import scipy.linalg
    atno, xyz = water_cluster(16, seed=16)
    B = O.Basis(atno, xyz)
    S, T, V = O.one_electron(B)
    nocc = int(sum(atno)) // 2
    lam, C = scipy.linalg.eigh(T + V, S, lower=True, driver="gvd")
    P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    P = 0.5 * (P + P.T)
    J, K, nq = O.jk_screened(B, P, thr=1e-12, nthreads=nthreads)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), atno=np.array(atno, dtype=np.int32), xyz=np.array(xyz),
                        P=P, J=np.asarray(J), K=np.asarray(K), quartets=np.array(nq), schwarz=np.array(1e-12))
    print(name, "n =", B.nbf, "screened quartets", nq, flush=True)


CASES = ("h2o_testgeom", "h2o_config0", "ch4", "water2_seed2", "water4_seed4", "water2_fd", "water16_jk")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cases", nargs="*", default=list(CASES), help="subset of %s" % (CASES,))
    want = set(ap.parse_args().cases)
    if "h2o_testgeom" in want:
        # reference test geometry (src/tests/test_gaussians_molecules.f90:164-168)
        run_case("h2o_testgeom", [8, 1, 1], [[0.0, 0.0, 0.0], [1.809, 0.0, 0.0], [-0.453549, 1.751221, 0.0]])
    if "h2o_config0" in want:
        # the benchmark monomer (O-H 0.9572 A, 104.52 deg)
        atno, xyz = water_monomer()
        run_case("h2o_config0", atno, xyz)
    if "ch4" in want:
        # CH4 geometry 1 (test_gaussians_molecules.f90:85-91)
        run_case("ch4", [6, 1, 1, 1, 1], [[0.0, 0.0, 0.0], [2.05, 0.0, 0.0], [-0.683292, 1.932773, 0.0],
                                          [-0.683292, -0.966387, -1.673831], [-0.683292, -0.966387, 1.673831]],
                 ndump=200)
    if "water2_seed2" in want:
        # seeded water dimer of the cluster generator (n = 50)
        atno, xyz = water_cluster(2, seed=2)
        run_case("water2_seed2", atno, xyz, Nscf=100, alpha=0.5, ndump=400)
    if "water4_seed4" in want:
        # four seeded waters (n = 100): the largest cluster under the reference's
        # n <= 120 guard (src/basis.f90:537)
        atno, xyz = water_cluster(4, seed=4)
        run_case("water4_seed4", atno, xyz, Nscf=300, alpha=0.5, ndump=400)
    if "water2_fd" in want:
        # reference energy differences for the gradient of the seeded dimer
        atno, xyz = water_cluster(2, seed=2)
        run_fd("water2_fd", atno, xyz)
    if "water16_jk" in want:
        run_jk16()


if __name__ == "__main__":
    main()
