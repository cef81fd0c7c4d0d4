"""Builds libtbfgpu.so (HIP for gfx950) in-tree with hipcc.

Usage: python -m tbfsolver_amd.build   (or tbfsolver_amd.build.build())
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtbfgpu.so")

SOURCES = ["common.cpp", "inputs.cpp", "fieldio.cpp", "tune.cpp", "pcg.cpp", "abi.cpp", "mg_kernels.hip", "sweep.hip", "fields.hip",
           "fps.hip", "vof.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TBFG_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-fopenmp",
          "-I" + os.path.join(HERE, "..", "include")]
# packed FP32 instructions stay off: hipcc forms them from ordinary float code
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    out = os.path.join(OBJ, src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), errs))
    return out


def build(verbose=False):
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fopenmp", "--offload-arch=" + ARCH, "-o", tmp] + objs + [
        "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib/llvm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_TEST = os.path.join(FORTRAN_DIR, "tbfgpu_test")


def build_fortran(verbose=False):
    """Compiles the iso_c_binding module and its test driver (needs flang)."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran binding not built")
        return None
    os.makedirs(OBJ, exist_ok=True)
    cmd = [FLANG, "-O2", "-J", OBJ, os.path.join(FORTRAN_DIR, "tbfgpu_mod.f90"),
           os.path.join(FORTRAN_DIR, "tbfgpu_test.f90"), "-o", FORTRAN_TEST, "-L" + HERE, "-ltbfgpu",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib/llvm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    if verbose:
        print("built", FORTRAN_TEST)
    return FORTRAN_TEST


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
