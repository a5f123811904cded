"""The Fortran iso_c_binding drop-in (hfsolver_amd/fortran/hfx_gpu.f90): a
Fortran program calls hfx_init + hfx_formG with gaussints-style arrays (1-based
istart, power(3,n), center(3,n)) and its G is checked against the compiled
reference's G(P) (tests/golden). Without a GPU it must stop with exit code 1
and the engine's message, as the reference's stop_error does."""
import os
import subprocess

import numpy as np
import pytest

from hfsolver_amd.basis import Basis

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "fortran", "hfx_fortran_check")


def _write_input(path, b, P):
    with open(path, "w") as f:
        f.write("%d %d %d\n" % (b.nbf, len(b.coef), len(b.atno)))
        for i in range(b.nbf):
            f.write("%d %d %d %d %d %.17e %.17e %.17e\n" % (b.nprim[i], b.istart[i] + 1, b.l[i], b.m[i], b.n[i],
                                                          b.x[i], b.y[i], b.z[i]))
        for c, a in zip(b.coef, b.alpha):
            f.write("%.17e %.17e\n" % (c, a))
        for z, r in zip(b.atno, b.xyz):
            f.write("%d %.17e %.17e %.17e\n" % (z, r[0], r[1], r[2]))
        for v in np.asarray(P).ravel(order="F"):
            f.write("%.17e\n" % v)


def _run(tmp_path, g):
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "fortran")], check=True)
    b = Basis(g["atno"], g["xyz"])
    inp = str(tmp_path / "in.txt")
    out = str(tmp_path / "out.txt")
    _write_input(inp, b, g["P"])
    r = subprocess.run([EXE, inp, out], capture_output=True, text=True)
    return b, r, out


@pytest.mark.gpu
def test_fortran_formG_matches_reference(golden, tmp_path):
    g = golden("water2_seed2")
    b, r, out = _run(tmp_path, g)
    assert r.returncode == 0, r.stdout + r.stderr
    G = np.loadtxt(out).reshape(b.nbf, b.nbf, order="F")
    assert np.abs(G - g["G"]).max() < 1e-11


def test_fortran_binding_stops_without_gpu(golden, tmp_path):
    g = golden("h2o_config0")
    b, r, out = _run(tmp_path, g)
    if r.returncode == 0:
        pytest.skip("a HIP device is present; covered by the gpu test")
    assert r.returncode == 1
    assert "hfx_create" in r.stdout and "device" in r.stdout
