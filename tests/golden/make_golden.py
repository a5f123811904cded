"""Write the golden fixtures from the compiled reference (run in the build
container only; /root/reference does not exist on the GPU box).

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

For each case it writes the input geometry, runs oracle/_ref/ref_driver (the
reference's own gaussints + doscf + get_nuclear_energy + kinetic_energy,
src/drivers.f90:29-41) with the working directory at the reference's test
folder (which holds p631ss.txt, src/basis.f90:210), and stores the outputs as
a compressed .npz (no pickles): energies, orbital energies, S, T, V, C, the
output density P, J(P), K(P), G(P) = J - K/2 from the reference's packed int2,
a strided sample of int2 with 1-based (i,j,k,l), and the per-iteration E0
printed by doscf (src/scf.f90:114).
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from hfsolver_amd.clusters import water_cluster, water_monomer  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REF_TESTS = "/root/reference/src/tests"


def ffloat(tok):
    """Fortran ES output drops the 'E' for 3-digit exponents (8.99-115)."""
    try:
        return float(tok)
    except ValueError:
        k = max(tok.rfind("-"), tok.rfind("+"))
        return float(tok[:k] + "E" + tok[k:])


def parse(path):
    d = {}
    lines = open(path).read().split("\n")
    i = 0
    while i < len(lines):
        t = lines[i].split()
        if not t:
            i += 1
            continue
        key, cnt = t[0], int(t[1])
        i += 1
        if cnt == -1:
            rows = [l.split() for l in lines[i:] if l.strip()]
            d["int2_idx"] = np.array([[int(r[0]), int(r[1]), int(r[2]), int(r[3])] for r in rows], dtype=np.int32)
            d["int2_val"] = np.array([ffloat(r[4]) for r in rows])
            break
        d[key] = np.array([ffloat(x) for x in lines[i:i + cnt]])
        i += cnt
    return d


def run_case(name, atno, xyz, Nscf=40, alpha=1.0, tolE=1e-10, tolP=1e-8, ndump=400):
    work = os.path.join(ROOT, "_work", "golden")
    os.makedirs(work, exist_ok=True)
    inp = os.path.join(work, name + ".in")
    out = os.path.join(work, name + ".out")
    with open(inp, "w") as f:
        f.write("%d\n" % len(atno))
        for z, r in zip(atno, xyz):
            f.write("%d %.17e %.17e %.17e\n" % (z, r[0], r[1], r[2]))
        f.write("%d %.17e %.17e %.17e\n%d\n" % (Nscf, alpha, tolE, tolP, ndump))
    env = dict(os.environ, MKL_NUM_THREADS="1", OMP_NUM_THREADS="1")
    log = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec %s %s %s" % (REF_BIN, inp, out)],
                         cwd=REF_TESTS, env=env, capture_output=True, text=True, check=True).stdout
    d = parse(out)
    n = int(d["n"][0])
    e_iter = []
    for line in log.split("\n"):
        t = line.split()
        if len(t) == 2 and t[0].isdigit():
            e_iter.append(float(t[1]))
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


def main():
    # reference test geometry (src/tests/test_gaussians_molecules.f90:164-168)
    run_case("h2o_testgeom", [8, 1, 1], [[0.0, 0.0, 0.0], [1.809, 0.0, 0.0], [-0.453549, 1.751221, 0.0]])
    # the benchmark monomer (O-H 0.9572 A, 104.52 deg)
    atno, xyz = water_monomer()
    run_case("h2o_config0", atno, xyz)
    # CH4 geometry 1 (test_gaussians_molecules.f90:85-91)
    run_case("ch4", [6, 1, 1, 1, 1], [[0.0, 0.0, 0.0], [2.05, 0.0, 0.0], [-0.683292, 1.932773, 0.0],
                                      [-0.683292, -0.966387, -1.673831], [-0.683292, -0.966387, 1.673831]],
             ndump=200)
    # seeded water dimer of the cluster generator (n = 50)
    atno, xyz = water_cluster(2, seed=2)
    run_case("water2_seed2", atno, xyz, Nscf=100, alpha=0.5, ndump=400)


if __name__ == "__main__":
    main()
