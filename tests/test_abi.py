"""The C-ABI library loads and exports every entry point include/mflbm.h declares (no GPU needed)."""
import ctypes as C
import os
import re

from mflbm import native

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    with open(os.path.join(REPO, "include", "mflbm.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(mflbm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = native.load()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(native.EXPORTED) == set(declared_symbols())


def test_create_without_gpu_fails_cleanly():
    lib = native.load()
    fn = lib.mflbm_create
    fn.restype = C.c_int32
    fn.argtypes = [C.POINTER(native.Params), C.POINTER(C.c_void_p)]
    p = native.Params()
    p.nx = p.ny = p.nz = 2   # rejected before any device call
    ctx = C.c_void_p()
    assert fn(C.byref(p), C.byref(ctx)) == 1
    assert b"at least" in lib.mflbm_create_error()
