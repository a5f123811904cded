"""Build the in-tree native library hfsolver_amd/libhfx.so (gfx950).

    python -m hfsolver_amd.build            # library only
    python -m hfsolver_amd.build --check    # also the host check harness of tests/

The kernels (csrc/hfx_kernels.hip, in parts, and csrc/hfx_lists.hip, the
quartet lists) are compiled for gfx950 only, with packed FP32 instructions
disabled; the host runtime (csrc/hfx_host.cpp) with OpenMP.
The Cython binding (_hfx.pyx over c_hfx.pxd) is compiled against the library
into hfsolver_amd/_hfx<EXT_SUFFIX>.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhfx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "--offload-arch=" + ARCH]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % os.path.basename(cmd[-1]))
    return r


def build_library(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    # the kernel source builds in parts (see its header), in parallel
    parts = {1: [], 2: ["-DHFX_REAL_T=double"], 3: ["-DHFX_REAL_T=float"], 4: ["-DHFX_REAL_T=double"],
             5: ["-DHFX_REAL_T=float"], 6: [], 7: []}
    objs = []
    jobs = []
    for part, extra in parts.items():
        o = os.path.join(BUILD, "hfx_kernels_%d.o" % part)
        objs.append(o)
        jobs.append(DEVICE_FLAGS + ["-DHFX_PART=%d" % part] + extra +
                    ["-c", os.path.join(CSRC, "hfx_kernels.hip"), "-o", o])
    objs.append(os.path.join(BUILD, "hfx_lists.o"))
    jobs.append(DEVICE_FLAGS + ["-c", os.path.join(CSRC, "hfx_lists.hip"), "-o", objs[-1]])
    objs.append(os.path.join(BUILD, "hfx_host.o"))
    jobs.append(HOST_FLAGS + ["-c", os.path.join(CSRC, "hfx_host.cpp"), "-o", objs[-1]])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(lambda j: run([HIPCC] + j), jobs))
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fopenmp", "--offload-arch=" + ARCH, "-o", tmp] + objs +
        ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib:/opt/rocm/llvm/lib"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_cython(verbose=True):
    """Cython -> C (_build/_hfx.c) -> hfsolver_amd/_hfx<EXT_SUFFIX>, linked to libhfx.so."""
    import sysconfig
    os.makedirs(BUILD, exist_ok=True)
    csrc = os.path.join(BUILD, "_hfx.c")
    run([sys.executable, "-m", "cython", "-3", "--module-name", "hfsolver_amd._hfx", "-I", HERE,
         "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "_hfx.pyx"), "-o", csrc])
    ext = os.path.join(HERE, "_hfx" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = ext + ".tmp"
    run(["gcc", "-O2", "-fPIC", "-shared", "-I" + sysconfig.get_paths()["include"], "-I" + os.path.join(ROOT, "include"),
         csrc, "-L" + HERE, "-lhfx", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, ext)
    if verbose:
        print("built", ext)
    return ext


def main(argv):
    build_library()
    build_cython()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
