"""The ISO_C_BINDING module compiles with flang and its calls reach the C ABI with the right layout
(host set-up entry point only; no GPU needed)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from mflbm import native
from mflbm.native import i8ptr

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
FORT = os.path.join(REPO, "mf-lbm_amd", "fortran")
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_fortran_binding_links_and_matches_python(tmp_path):
    lib_dir = os.path.dirname(native.LIB)
    subprocess.run([FLANG, "-c", os.path.join(FORT, "mflbm_c_binding.f90"), "-module-dir", str(tmp_path),
                    "-o", str(tmp_path / "b.o")], check=True)
    exe = tmp_path / "binding_check"
    subprocess.run([FLANG, os.path.join(FORT, "binding_check.f90"), str(tmp_path / "b.o"), f"-I{tmp_path}",
                    f"-L{lib_dir}", "-lmflbm", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout
    rc, chk = [int(t.split("=")[1]) for t in out.split()]
    assert rc == 0
    # same call from Python
    lib = native.load()
    p = native.Params()
    p.nx, p.ny, p.nz = 8, 6, 5
    p.jper, p.kper = 0, 1
    p.wall_x_min = p.wall_x_max = p.wall_y_min = p.wall_y_max = 1
    wg = np.zeros((5, 6, 8), np.int8)
    wg[1, 2, 3] = 1
    wh = np.zeros((9, 10, 12), np.int8)
    assert lib.mflbm_setup_walls(C.byref(p), i8ptr(wg), i8ptr(wh)) == 0
    kk, jj, ii = np.meshgrid(np.arange(-1, 5 + 3), np.arange(-1, 6 + 3), np.arange(-1, 8 + 3), indexing="ij")
    s = int((wh.astype(np.int64) * (ii + 3 * jj + 7 * kk + 100)).sum())
    assert s == chk
