"""Seeded synthetic water clusters (geometry inputs for tests and benchmarks).

Monomer: O at the origin, H1 at (R, 0, 0), H2 at (R cos t, R sin t, 0) with
R = 0.9572 Angstrom, t = 104.52 degrees, converted to bohr with the reference's
bohr2ang = 0.529177249 (hfsolver src/constants.f90:43-44).

Cluster of N waters: sites of a cubic grid of side s = ceil(N^(1/3)) with
3.0 Angstrom spacing (the first N sites of a seeded permutation of the s^3
sites), each O jittered uniformly in [-0.3, 0.3] Angstrom per coordinate (a
site's jitter is re-drawn while any O-O distance is below 2.5 Angstrom), each
molecule rotated by a uniform random rotation (Shoemake's quaternion from three
uniforms). The generator is numpy's PCG64 seeded with `seed`.
"""
import math

import numpy as np

BOHR2ANG = 0.529177249
ANG2BOHR = 1.0 / BOHR2ANG
OH_ANG = 0.9572
HOH_DEG = 104.52


def water_monomer():
    """(atno, xyz[3,3] bohr) of the rigid water of the benchmark configurations."""
    R = OH_ANG * ANG2BOHR
    t = math.radians(HOH_DEG)
    xyz = np.array([[0.0, 0.0, 0.0], [R, 0.0, 0.0], [R * math.cos(t), R * math.sin(t), 0.0]])
    return [8, 1, 1], xyz


def _quat_rotation(u1, u2, u3):
    a = math.sqrt(1 - u1) * math.sin(2 * math.pi * u2)
    b = math.sqrt(1 - u1) * math.cos(2 * math.pi * u2)
    c = math.sqrt(u1) * math.sin(2 * math.pi * u3)
    d = math.sqrt(u1) * math.cos(2 * math.pi * u3)
    w, x, y, z = d, a, b, c
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(nwater, seed=0, spacing=3.0, jitter=0.3, min_oo=2.5):
    """(atno list, xyz[3*nwater, 3] in bohr) of a seeded water cluster."""
    if nwater < 1:
        raise ValueError("nwater must be >= 1")
    rng = np.random.Generator(np.random.PCG64(seed))
    s = 1
    while s ** 3 < nwater:
        s += 1
    sites = np.array([(i, j, k) for i in range(s) for j in range(s) for k in range(s)], dtype=float)
    order = rng.permutation(len(sites))[:nwater]
    centres = sites[order] * spacing
    oxy = []
    for c in centres:
        for _ in range(1000):
            o = c + rng.uniform(-jitter, jitter, size=3)
            if all(np.linalg.norm(o - p) >= min_oo for p in oxy):
                break
        else:
            raise RuntimeError("could not place water %d" % len(oxy))
        oxy.append(o)
    _, mono = water_monomer()
    atno, xyz = [], []
    for o in oxy:
        Rm = _quat_rotation(*rng.uniform(0.0, 1.0, size=3))
        for z, r in zip([8, 1, 1], mono):
            atno.append(z)
            xyz.append(o * ANG2BOHR + Rm @ r)
    return atno, np.array(xyz)
