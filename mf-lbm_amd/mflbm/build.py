"""Build the native libraries in-tree.

* ``mflbm/lib/libmflbm.so``: the gfx950 device code of the two-phase and single-phase paths and the C ABI
  (hipcc, FP contraction off so the FP64
  arithmetic rounds exactly as written; packed-FP32 instructions disabled).
* ``oracle/liboracle.so``: the CPU checker (g++, OpenMP), built only when its sources are present.
* ``mf-lbm_amd/fortran/bin/hotpath_run``: the Fortran main program that drives the hot path through the
  ISO_C_BINDING module (flang), when flang is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                 # mf-lbm_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG, "lib", "libmflbm.so")
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MFLBM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_device(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(CSRC, "mflbm_device.hip"), os.path.join(CSRC, "sp_device.hip"),
            os.path.join(CSRC, "sp_batch.hip"), os.path.join(CSRC, "nn_kernels.hip"),
            os.path.join(CSRC, "geometry_device.hip"), os.path.join(CSRC, "host_setup.cpp")]
    tmp = LIB + ".tmp"
    out = _run([HIPCC, *HIP_FLAGS, f"-I{CSRC}", f"-I{INCLUDE}", *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print(out)
    return LIB


def build_oracle(verbose: bool = False) -> str | None:
    src = os.path.join(ORACLE_DIR, "oracle.cpp")
    if not os.path.exists(src):
        return None
    srcs = [src, os.path.join(ORACLE_DIR, "oracle_sp.cpp")]
    cxx = shutil.which("g++") or "g++"
    tmp = ORACLE_LIB + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
          *srcs, "-o", tmp])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


FORTRAN_DIR = os.path.join(ROOT, "fortran")
FORTRAN_BIN = os.path.join(FORTRAN_DIR, "bin", "hotpath_run")


def build_fortran(verbose: bool = False) -> str | None:
    """The Fortran driver, linked against the in-tree libmflbm.so (found at run time relative to the binary)."""
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(flang):
        return None
    out_dir = os.path.dirname(FORTRAN_BIN)
    mod_dir = os.path.join(out_dir, "mod")
    os.makedirs(mod_dir, exist_ok=True)
    obj = os.path.join(mod_dir, "mflbm_c_binding.o")
    _run([flang, "-O2", "-c", os.path.join(FORTRAN_DIR, "mflbm_c_binding.f90"), "-module-dir", mod_dir, "-o", obj])
    tmp = FORTRAN_BIN + ".tmp"
    _run([flang, "-O2", os.path.join(FORTRAN_DIR, "hotpath_run.f90"), obj, f"-I{mod_dir}",
          f"-L{os.path.dirname(LIB)}", "-lmflbm", "-Wl,-rpath,$ORIGIN/../../mflbm/lib", "-o", tmp])
    os.replace(tmp, FORTRAN_BIN)
    return FORTRAN_BIN


def build_all(verbose: bool = False) -> None:
    build_device(verbose)
    build_oracle(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
