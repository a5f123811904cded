"""Loader of the CPU check harness tests/csrc/liberi_check.so (the engine's
Rys templates instantiated for the host). Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "eri_check.cpp")
LIB = os.path.join(HERE, "csrc", "liberi_check.so")


SAN_SRC = os.path.join(HERE, "csrc", "san_check.cpp")
SAN_EXE = os.path.join(HERE, "csrc", "san_check")


def build():
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", SRC, "-o", LIB], check=True)
    build_sanitized()


def build_sanitized():
    """The harness and its driver (san_check.cpp) as one program built with
    AddressSanitizer and UndefinedBehaviorSanitizer (host code only)."""
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", SRC, SAN_SRC, "-o", SAN_EXE], check=True)


ROOT = os.path.dirname(HERE)
HOST_SAN_SRC = os.path.join(HERE, "csrc", "host_san.cpp")
HOST_SAN_EXE = os.path.join(HERE, "csrc", "host_san")
HOST_SAN_SUPP = os.path.join(HERE, "csrc", "lsan_rocm.supp")


def build_host_sanitized():
    """The engine's host runtime (hfsolver_amd/csrc/hfx_host.cpp) and the driver
    host_san.cpp built with AddressSanitizer and UndefinedBehaviorSanitizer (host
    code only, sanitizer runtimes linked statically), linked with the gfx950
    kernel objects of the library build (hfsolver_amd/_build). Runs on a GPU."""
    kobj = [os.path.join(ROOT, "hfsolver_amd", "_build", "hfx_kernels_%d.o" % p) for p in range(1, 8)]
    kobj.append(os.path.join(ROOT, "hfsolver_amd", "_build", "hfx_lists.o"))
    missing = [o for o in kobj if not os.path.exists(o)]
    if missing:
        raise RuntimeError("kernel objects missing (python -m hfsolver_amd.build first): %s" % missing[0])
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
    host_o = os.path.join(HERE, "csrc", "host_san_hfx_host.o")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"] + san +
                   ["-c", os.path.join(ROOT, "hfsolver_amd", "csrc", "hfx_host.cpp"), "-o", host_o], check=True)
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-I" + os.path.join(ROOT, "include")] + san +
                   ["-static-libasan", "-static-libubsan", HOST_SAN_SRC, host_o] + kobj +
                   ["-fopenmp", "-L/opt/rocm/lib", "-lamdhip64", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib",
                    "-o", HOST_SAN_EXE], check=True)


def load():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        build()
    return ctypes.CDLL(LIB)


def shells_of(B):
    """Shells of a per-function basis (L, first function, centre, exponents,
    shell coefficients = first component's, per-component scale)."""
    out = []
    i = 0
    while i < B.nbf:
        L = int(B.l[i] + B.m[i] + B.n[i])
        nc = (L + 1) * (L + 2) // 2
        k = int(B.nprim[i])
        s = int(B.istart[i])
        c0 = B.coef[s:s + k].copy()
        sc = np.array([B.coef[B.istart[i + c]] / c0[0] for c in range(nc)])
        out.append(dict(L=L, f0=i, cen=np.array([B.x[i], B.y[i], B.z[i]]), a=B.alpha[s:s + k].copy(), c=c0, sc=sc))
        i += nc
    return out


def pair_class(la, lb):
    return 0 if la == 0 else ((1 if lb == 0 else 2) if la == 1 else 3 + lb)


def quartet(lib, A, B, C, D, fp32=False):
    """Block (ab|cd) of shells oriented l_A >= l_B, l_C >= l_D, class(AB) >= class(CD)."""
    nc = lambda s: (s["L"] + 1) * (s["L"] + 2) // 2
    out = np.zeros(nc(A) * nc(B) * nc(C) * nc(D))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    args = [A["L"], B["L"], C["L"], D["L"]] + [p(s["cen"]) for s in (A, B, C, D)]
    for s in (A, B, C, D):
        args += [len(s["a"]), p(s["a"]), p(s["c"])]
    rc = lib.chk_eri_quartet(*args, 1 if fp32 else 0, p(out))
    if rc != 0:
        raise ValueError("bad quartet orientation")
    blk = out.reshape(nc(A), nc(B), nc(C), nc(D))
    return blk * np.einsum("a,b,c,d->abcd", A["sc"], B["sc"], C["sc"], D["sc"])
