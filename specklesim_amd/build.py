"""In-tree build of libspeckle_gpu.so (hipcc, gfx950).

The shared library lands next to this file so that it travels with the
repository; ``python -m specklesim_amd.build`` rebuilds it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libspeckle_gpu.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIP_SOURCES = ["spk_kernels.hip", "spk_api.hip", "spk_quad.hip", "spk_batch.hip", "spk_tree.hip"]
CXX_SOURCES = ["spk_host.cpp", "spk_quad_host.cpp", "spk_ensemble.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPK_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 are generated.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-I", os.path.join(HERE, "..", "include"), "-I", "/opt/rocm/include",
             "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return p.stdout + p.stderr


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(BUILD_DIR, s + ".o")
        jobs.append((obj, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", obj]))
    for s in CXX_SOURCES:
        obj = os.path.join(BUILD_DIR, s + ".o")
        jobs.append((obj, [shutil.which("g++") or "g++", *CXX_FLAGS, "-c", os.path.join(CSRC, s),
                           "-o", obj]))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[1]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp,
          *[j[0] for j in jobs], "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-lpthread",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_EXE = os.path.join(HERE, "spk_fortran_check")


def build_fortran() -> str | None:
    """The Fortran ISO_C_BINDING module and its driver program (a Fortran
    host calling the library the way the reference's solver does).  Skipped
    when no Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(BUILD_DIR, "fortran")
    os.makedirs(mod, exist_ok=True)
    obj = os.path.join(mod, "spk_gpu_binding.o")
    _run([FLANG, "-O2", "-module-dir", mod, "-c",
          os.path.join(FORTRAN_DIR, "spk_gpu_binding.f90"), "-o", obj])
    tmp = FORTRAN_EXE + ".tmp"
    _run([FLANG, "-O2", "-module-dir", mod, "-I", mod,
          os.path.join(FORTRAN_DIR, "spk_fortran_check.f90"), obj, "-L", HERE, "-lspeckle_gpu",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, FORTRAN_EXE)
    return FORTRAN_EXE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
