"""The ISO_C_BINDING module compiles with flang and its calls reach the C ABI with the right layout
(host set-up entry point only; no GPU needed)."""
import ctypes as C
import os
import re
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
    vals = dict(re.findall(r"(\w+)=\s*(\S+)", out))
    rc, chk = int(vals["rc"]), int(vals["checksum"])
    sp_rc, sp_sum = int(vals["sp_rc"]), float(vals["sp_sum"])
    assert rc == 0 and sp_rc == 0
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

    # single-phase init through the same interface
    from mflbm.native import SPParams, dptr, ptr_array
    q = SPParams()
    q.nx, q.ny, q.nz, q.kper = 8, 6, 5, 1
    q.wall_x_min = q.wall_x_max = q.wall_y_min = q.wall_y_max = 1
    q.mrt_preset, q.la_nu = 1, 0.1
    f = np.zeros((19, 7, 8, 10))
    u = np.zeros((7, 8, 10))
    u[:, :, :] = 0.01 * np.arange(10)[None, None, :]
    assert lib.mflbm_sp_init_fields(C.byref(q), None, dptr(u), None, None, ptr_array(list(f)), None) == 0
    t = 0.0
    for m in range(19):
        t = t + f[m].sum() * (m + 1)
    assert sp_sum == pytest.approx(t, rel=1e-14)
