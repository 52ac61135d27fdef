"""ctypes view of the C ABI declared in include/mflbm.h.

The library is the in-tree build ``mflbm/lib/libmflbm.so`` (see mflbm.build).  There is no fallback: if the
library is missing or fails to load, importing the solver raises, so a run can never silently use a
different implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .build import LIB

c_int = C.c_int32
c_dbl = C.c_double
P_dbl = C.POINTER(C.c_double)
P_int = C.POINTER(C.c_int32)
P_i8 = C.POINTER(C.c_int8)


class Params(C.Structure):
    """mflbm_params"""
    _fields_ = [(n, c_int) for n in (
        "nx", "ny", "nz", "jper", "kper", "wall_x_min", "wall_x_max", "wall_y_min", "wall_y_max",
        "wall_z_min", "wall_z_max", "inlet_bc", "outlet_bc", "porous_plate_cmd", "z_porous_plate",
        "n_exclude_inlet", "n_exclude_outlet", "precision")] + [(n, c_dbl) for n in (
        "la_nui1", "la_nui2", "gamma", "beta", "force_z", "sa_inject", "phi_inlet", "rho_in", "rho_out",
        "uin_avg", "relaxation")]


class SPParams(C.Structure):
    """mflbm_sp_params (single-phase solver)"""
    _fields_ = [(n, c_int) for n in (
        "nx", "ny", "nz", "jper", "kper", "wall_x_min", "wall_x_max", "wall_y_min", "wall_y_max",
        "wall_z_min", "wall_z_max", "inlet_bc", "outlet_bc", "n_exclude_inlet", "n_exclude_outlet",
        "mrt_preset", "precision")] + [(n, c_dbl) for n in (
        "la_nu", "force_z", "rho_in", "rho_out", "uin_avg", "relaxation")]


class SPBParams(C.Structure):
    """mflbm_spb_params (batched single-phase runs)"""
    _fields_ = [(n, c_int) for n in ("precision", "mrt_preset", "ntime_monitor", "tile_rows", "block", "rest_state")] + \
        [(n, c_dbl) for n in ("la_nu", "force_z", "tol")]


class MonitorSums(C.Structure):
    """mflbm_monitor_sums"""
    _fields_ = [(n, c_dbl) for n in ("umax2", "usq1", "usq2", "vol1_full", "vol2_full", "dphi_max",
                                     "pre_w", "pre_nw")] + [(n, C.c_int64) for n in ("n_w", "n_nw", "outlet_phase1")]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_dbl)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(P_int)


def i8ptr(a: np.ndarray):
    assert a.dtype == np.int8 and a.flags.c_contiguous
    return a.ctypes.data_as(P_i8)


def ptr_array(arrs):
    """19 population pointers for the (const) double* const* arguments."""
    return (P_dbl * len(arrs))(*[dptr(a) for a in arrs])


_ENTRY_POINTS = {
    # name: (restype, argtypes)
    "create": (c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "destroy": (None, [C.c_void_p]),
    "last_error": (C.c_char_p, [C.c_void_p]),
    "set_scalars": (c_int, [C.c_void_p, c_dbl, c_dbl, c_dbl, c_dbl]),
    "upload_geometry": (c_int, [C.c_void_p, P_i8, c_int, P_int, P_int, P_int, P_dbl, c_int, P_int, P_dbl, P_dbl, P_dbl]),
    "upload_state": (c_int, [C.c_void_p, C.POINTER(P_dbl), C.POINTER(P_dbl), P_dbl, P_dbl, P_dbl, P_dbl]),
    "download_state": (c_int, [C.c_void_p, C.POINTER(P_dbl), C.POINTER(P_dbl), P_dbl, P_dbl, P_dbl, P_dbl]),
    "init_state_device": (c_int, [C.c_void_p, c_int, c_dbl]),
    "step": (c_int, [C.c_void_p, c_int]),
    "run": (c_int, [C.c_void_p, c_int, c_int]),
    "color_gradient": (c_int, [C.c_void_p]),
    "compute_macro": (c_int, [C.c_void_p]),
    "monitor": (c_int, [C.c_void_p, P_dbl, C.POINTER(MonitorSums), c_int]),
    "download_macro": (c_int, [C.c_void_p, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]),
    "download_gradient": (c_int, [C.c_void_p, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]),
    "synchronize": (c_int, [C.c_void_p]),
    "timing": (c_int, [C.c_void_p, c_int, P_dbl, C.POINTER(C.c_int64)]),
    "timing_enable": (c_int, [C.c_void_p, c_int]),
}

# single-phase context entry points: name -> (restype, argtypes), bound as prefix + "sp_" + name
_SP_ENTRY_POINTS = {
    "create": (c_int, [C.POINTER(SPParams), C.POINTER(C.c_void_p)]),
    "destroy": (None, [C.c_void_p]),
    "last_error": (C.c_char_p, [C.c_void_p]),
    "set_scalars": (c_int, [C.c_void_p, c_dbl, c_dbl, c_dbl, c_dbl]),
    "upload_geometry": (c_int, [C.c_void_p, P_i8, P_dbl]),
    "upload_state": (c_int, [C.c_void_p, C.POINTER(P_dbl), P_dbl]),
    "download_state": (c_int, [C.c_void_p, C.POINTER(P_dbl), P_dbl]),
    "step": (c_int, [C.c_void_p, c_int]),
    "run": (c_int, [C.c_void_p, c_int, c_int]),
    "compute_macro": (c_int, [C.c_void_p]),
    "monitor": (c_int, [C.c_void_p, P_dbl, P_dbl]),
    "download_macro": (c_int, [C.c_void_p, P_dbl, P_dbl, P_dbl, P_dbl]),
    "synchronize": (c_int, [C.c_void_p]),
    "timing": (c_int, [C.c_void_p, c_int, P_dbl, C.POINTER(C.c_int64)]),
    "timing_enable": (c_int, [C.c_void_p, c_int]),
    "set_equilibrium": (c_int, [C.c_void_p, P_dbl, P_dbl, P_dbl, P_dbl]),
}

# batched single-phase entry points, bound as prefix + "spb_" + name
_SPB_ENTRY_POINTS = {
    "create": (c_int, [C.POINTER(SPBParams), C.POINTER(C.c_void_p)]),
    "destroy": (None, [C.c_void_p]),
    "last_error": (C.c_char_p, [C.c_void_p]),
    "add_sample": (c_int, [C.c_void_p, c_int, c_int, c_int, c_int, P_i8]),
    "commit": (c_int, [C.c_void_p]),
    "info": (c_int, [C.c_void_p, P_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "sample_fluid": (c_int, [C.c_void_p, c_int, C.POINTER(C.c_int64)]),
    "upload_state": (c_int, [C.c_void_p, c_int, C.POINTER(P_dbl)]),
    "download_state": (c_int, [C.c_void_p, c_int, C.POINTER(P_dbl)]),
    "run": (c_int, [C.c_void_p, c_int, c_int]),
    "run_steady": (c_int, [C.c_void_p, c_int, P_int]),
    "results": (c_int, [C.c_void_p, P_int, P_dbl, P_int]),
    "reset_monitor": (c_int, [C.c_void_p]),
    "synchronize": (c_int, [C.c_void_p]),
    "download_macro": (c_int, [C.c_void_p, c_int, P_dbl, P_dbl, P_dbl, P_dbl]),
}

# warm-start network kernels (device pointers as void*, stream as void*)
_VP = C.c_void_p
_NN_ENTRY_POINTS = {
    "mflbm_nn_weights_to_bf16": (c_int, [_VP, _VP, _VP, c_int, c_int, _VP]),
    "mflbm_nn_conv3d_fwd": (c_int, [_VP, _VP, _VP, _VP, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _VP]),
    "mflbm_nn_conv3d_dw": (c_int, [_VP, _VP, _VP, _VP, c_int, c_int, c_int, c_int, c_int, c_int, _VP]),
    "mflbm_nn_conv3d_dw_work": (C.c_int64, [c_int, c_int, c_int, c_int, c_int, c_int]),
    "mflbm_nn_conv3d_dw_ordered": (c_int, [_VP, _VP, _VP, _VP, _VP, C.c_int64, c_int, c_int, c_int, c_int, c_int, c_int,
                                           _VP]),
    "mflbm_nn_conv3d_dw_scratch_floats": (C.c_int64, []),
    "mflbm_nn_conv3d_dw_ordered_noalloc": (c_int, [_VP, _VP, _VP, _VP, c_int, c_int, c_int, c_int, c_int, c_int, _VP]),
    "mflbm_nn_relu_bwd": (c_int, [_VP, _VP, C.c_int64, _VP]),
    "mflbm_nn_pool2": (c_int, [_VP, _VP, c_int, c_int, c_int, c_int, c_int, c_int, _VP]),
    "mflbm_nn_up2": (c_int, [_VP, _VP, c_int, c_int, c_int, c_int, c_int, c_int, _VP]),
    "mflbm_nn_add": (c_int, [_VP, _VP, _VP, C.c_int64, _VP]),
    "mflbm_nn_mse": (c_int, [_VP, _VP, _VP, _VP, C.c_int64, c_int, c_int, C.c_float, _VP]),
    "mflbm_nn_adam": (c_int, [_VP, _VP, _VP, _VP, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, c_int, _VP]),
    "mflbm_nn_mfma_probe": (c_int, [_VP, _VP, _VP, _VP]),
    "mflbm_nn_distance_features": (c_int, [_VP, _VP, _VP, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _VP]),
}

_HOST_ENTRY_POINTS = {
    "mflbm_create_error": (C.c_char_p, []),
    "mflbm_setup_walls": (c_int, [C.POINTER(Params), P_i8, P_i8]),
    "mflbm_geometry_new": (c_int, [C.POINTER(Params), P_i8, c_dbl, C.POINTER(C.c_void_p)]),
    "mflbm_geometry_new_device": (c_int, [C.POINTER(Params), P_i8, c_dbl, C.POINTER(C.c_void_p)]),
    "mflbm_geometry_counts": (c_int, [C.c_void_p, P_int, P_int]),
    "mflbm_geometry_export": (c_int, [C.c_void_p, P_int, P_int, P_int, P_dbl, P_int, P_dbl, P_dbl]),
    "mflbm_geometry_free": (None, [C.c_void_p]),
    "mflbm_init_fields": (c_int, [C.POINTER(Params), c_int, c_dbl, P_dbl, P_dbl, P_dbl, C.POINTER(P_dbl),
                                  C.POINTER(P_dbl), P_dbl, P_dbl, P_dbl, P_dbl]),
    "mflbm_inlet_profile": (c_int, [C.POINTER(Params), c_dbl, c_dbl, c_dbl, P_dbl]),
    "mflbm_sp_init_fields": (c_int, [C.POINTER(SPParams), P_dbl, P_dbl, P_dbl, P_dbl, C.POINTER(P_dbl), P_dbl]),
}

# every symbol include/mflbm.h declares
EXPORTED = (["mflbm_" + n for n in _ENTRY_POINTS] + ["mflbm_sp_" + n for n in _SP_ENTRY_POINTS] +
            ["mflbm_spb_" + n for n in _SPB_ENTRY_POINTS] + list(_HOST_ENTRY_POINTS) + list(_NN_ENTRY_POINTS))


def bind(lib: C.CDLL, prefix: str, names=None, single_phase: bool = False, batch: bool = False):
    """Attach restype/argtypes to ``prefix + name`` for the context entry points (two-phase, with
    ``single_phase`` the ``prefix + "sp_" + name`` ones, with ``batch`` the ``prefix + "spb_" + name`` ones);
    returns a dict keyed by ``name``."""
    out = {}
    table = _SPB_ENTRY_POINTS if batch else (_SP_ENTRY_POINTS if single_phase else _ENTRY_POINTS)
    if batch:
        prefix = prefix + "spb_"
    elif single_phase:
        prefix = prefix + "sp_"
    for name, (res, args) in table.items():
        if names is not None and name not in names:
            continue
        if not hasattr(lib, prefix + name):
            continue
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
        out[name] = fn
    return out


_lib = None


def load() -> C.CDLL:
    """Load the in-tree HIP library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                               "or `python -m mflbm.build`")
        # RTLD_GLOBAL: the library's HIP and HSA runtimes (ROCm's) join the global symbol scope, so a torch
        # imported later binds its own HIP runtime to the same HSA runtime instead of starting a second one,
        # which could not take the GPU's virtual address space this process already holds
        lib = C.CDLL(LIB, mode=C.RTLD_GLOBAL)
        for name, (res, args) in list(_HOST_ENTRY_POINTS.items()) + list(_NN_ENTRY_POINTS.items()):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
