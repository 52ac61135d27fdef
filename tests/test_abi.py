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


def test_allocation_free_weight_gradient_refuses_oversize_before_any_device_call():
    """The size check of mflbm_nn_conv3d_dw_ordered_noalloc runs on the host before the scratch lookup or any
    launch, so it holds without a GPU (the pointers are never dereferenced)."""
    lib = native.load()
    assert lib.mflbm_nn_conv3d_dw_scratch_floats() == 4 << 20
    shape, cin, cout = (1, 32, 32, 160), 64, 64
    assert lib.mflbm_nn_conv3d_dw_work(*shape, cin, cout) > lib.mflbm_nn_conv3d_dw_scratch_floats()
    fake = C.c_void_p(4096)
    assert lib.mflbm_nn_conv3d_dw_ordered_noalloc(fake, fake, fake, None, *shape, cin, cout, None) == 1
    # and the argument checks of the workspace form
    assert lib.mflbm_nn_conv3d_dw_ordered_noalloc(None, fake, fake, None, 1, 8, 8, 8, 8, 8, None) == 1
    assert lib.mflbm_nn_conv3d_dw_ordered_noalloc(fake, fake, fake, None, 1, 8, 8, 12, 8, 8, None) == 1
