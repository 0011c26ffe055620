"""Build recipe of libflowmol_hip.so (gfx950) -- hipcc directly, in-tree.

python -m flowmol_amd.build [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libflowmol_hip.so")
SOURCES = ["fm_api.hip", "fm_kernels.hip", "fm_brick.hip", "fm_batch.hip"]
ARCH = os.environ.get("FLOWMOL_ARCH", "gfx950")

CFLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    # exact FP64 operation order (no fused multiply-add unless a kernel opts in)
    "-ffp-contract=off",
    # hardware FP64 atomic add on coarse-grained device memory
    "-munsafe-fp-atomics",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_BIN = os.path.join(FORTRAN_DIR, "bin")


def amdflang():
    for c in (os.environ.get("FC_AMD"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build_fortran(verbose=True):
    """The iso_c_binding interface module and the Fortran time-loop driver (amdflang)."""
    fc = amdflang()
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran interface not built")
        return None
    os.makedirs(FORTRAN_BIN, exist_ok=True)
    mod_obj = os.path.join(FORTRAN_BIN, "flowmol_hip_c.o")
    cmds = [
        [fc, "-O2", "-c", os.path.join(FORTRAN_DIR, "flowmol_hip_c.f90"), "-J", FORTRAN_BIN, "-o", mod_obj],
        [fc, "-O2", os.path.join(FORTRAN_DIR, "md_vv_driver.f90"), mod_obj, "-J", FORTRAN_BIN, "-L", HERE,
         "-lflowmol_hip", "-Wl,-rpath,$ORIGIN/../..", "-o", os.path.join(FORTRAN_BIN, "md_vv_driver")],
    ]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"Fortran build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return os.path.join(FORTRAN_BIN, "md_vv_driver")


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs or len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs)
    sys.exit(0)
