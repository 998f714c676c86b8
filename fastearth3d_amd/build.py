"""Build the HIP library fastearth3d_amd/libvilma_hip.so for gfx950.

Each translation unit of csrc/ is compiled with hipcc (in parallel), then
linked into one shared library exporting the C-ABI of include/vilma_hip.h.
Usage: python -m fastearth3d_amd.build  (or build() from Python).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvilma_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["vh_setup.hip", "vh_drift.hip", "vh_advance.hip", "vh_sht.hip", "vh_sle.hip", "vh_rot.hip",
           "vh_tensor.hip", "vh_toroidal.hip", "vh_adjoint.hip", "vh_remap.hip", "vh_api.hip", "vh_ensemble.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran", "vilma.f90")
DRIVER_SRC = os.path.join(HERE, "..", "tests", "fortran", "drive.f90")
DRIVER = os.path.join(OBJ, "vilma_drive")


def build_fortran(verbose=False):
    """Compile the Fortran host module `vilma` (object + .mod in _build/) and
    link the test driver against libvilma_hip.so (rpath relative to itself).
    Skipped, with a message, where no flang is installed."""
    if not os.path.exists(FLANG):
        if verbose:
            print("flang not found: Fortran module not built")
        return None
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, "vilma_mod.o")
    cmds = [[FLANG, "-O2", "-fPIC", "-module-dir", OBJ, "-c", FORTRAN, "-o", obj]]
    if os.path.exists(DRIVER_SRC):
        cmds.append([FLANG, "-O2", "-I", OBJ, DRIVER_SRC, obj, "-L", HERE, "-lvilma_hip",
                     "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER])
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("flang failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose:
        print("built", obj, DRIVER)
    return DRIVER


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
