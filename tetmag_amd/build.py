"""Build recipe for libtetmag_hip.so (gfx950) and the facade test driver.

Compiles every HIP/C++ source in tetmag_amd/csrc with hipcc into
tetmag_amd/libtetmag_hip.so.  Run as ``python -m tetmag_amd.build`` or call
:func:`build`.  No GPU is needed to build.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtetmag_hip.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("TETMAG_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}", f"-I{CSRC}"]

SOURCES = [
    "tm_sparse.hip",
    "tm_cg.hip",
    "tm_amg.hip",
    "tm_bem.hip",
    "tm_field.hip",
    "tm_batch.hip",
    "tm_adjoint.hip",
    "tm_integrate.hip",
    "tm_setup.cpp",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-fopenmp", "-x", "c++", "-c", path, "-o", obj]
    out = _run(cmd)
    return obj, out


def build(verbose: bool = False) -> str:
    """Compile all sources and link the shared library; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for o, out in results:
            if out.strip():
                print(o, out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fopenmp", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/llvm/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
