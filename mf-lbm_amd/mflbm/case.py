"""Inputs of one run, prepared on the host the way the reference driver prepares them.

``Case(control, sample)`` takes the control-file values and the wall sample and produces, through the host
set-up entry points of the C ABI:

* ``params``: the per-run constants (initialization_basic_multi, Init_multiphase.F90:5-191),
* ``walls_h2``: the local wall array with its 2-layer halo (set_walls, Misc.F90:6-210),
* the solid/fluid boundary-node lists with wall normals (geometry_preprocessing_new,
  Geometry_preprocessing.F90:9-512),
* ``w_in``: the inlet velocity profile (Init_multiphase.F90:194-236, Misc.F90:625-665),
* ``initial_state()``: the fields of a new run (initialization_new_multi, Init_multiphase.F90:243-470).

Array shapes follow the reference: populations (nz+2, ny+2, nx+2), phi (nz+8, ny+8, nx+8), walls
(nz+4, ny+4, nx+4), planes (ny+2, nx+2), indexed [k, j, i].
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Dict, Optional

import numpy as np

from . import io, native
from .native import dptr, i8ptr, iptr


@dataclasses.dataclass
class State:
    f: np.ndarray                 # (19, nz+2, ny+2, nx+2)
    g: np.ndarray
    phi: np.ndarray               # (nz+8, ny+8, nx+8)
    f_convec: np.ndarray          # (19, ny+2, nx+2)
    g_convec: np.ndarray
    phi_convec: np.ndarray        # (ny+2, nx+2)

    @staticmethod
    def empty(nx: int, ny: int, nz: int) -> "State":
        return State(np.zeros((19, nz + 2, ny + 2, nx + 2)), np.zeros((19, nz + 2, ny + 2, nx + 2)),
                     np.zeros((nz + 8, ny + 8, nx + 8)), np.zeros((19, ny + 2, nx + 2)),
                     np.zeros((19, ny + 2, nx + 2)), np.zeros((ny + 2, nx + 2)))

    @staticmethod
    def from_checkpoint(ck: io.Checkpoint) -> "State":
        nz2, ny2, nx2 = ck.f.shape[1:]
        z = lambda *s: np.zeros(s)
        return State(np.ascontiguousarray(ck.f), np.ascontiguousarray(ck.g), np.ascontiguousarray(ck.phi),
                     ck.f_convec if ck.f_convec is not None else z(19, ny2, nx2),
                     ck.g_convec if ck.g_convec is not None else z(19, ny2, nx2),
                     ck.phi_convec if ck.phi_convec is not None else z(ny2, nx2))


@dataclasses.dataclass
class BoundaryLists:
    s_ijk: np.ndarray    # (ns, 3) int32
    s_nfl: np.ndarray    # (ns,)
    s_nbr: np.ndarray    # (ns, 18)
    s_w: np.ndarray      # (ns,)
    f_ijk: np.ndarray    # (nf, 3)
    f_nw: np.ndarray     # (nf, 3)
    f_theta: np.ndarray  # (nf,)


PRECISION = {"fp64": 0, "fp32": 1, "mixed": 2}


def make_params(rp: io.RunParams, precision: str = "fp64") -> native.Params:
    """precision "fp64": the reference arithmetic; "fp32": populations stored in FP32 (f + g - REST and
    f - g) with FP32 collision and recolouring; "mixed": the same storage collided and recoloured in FP64;
    everything else FP64 in every mode."""
    if precision not in PRECISION:
        raise ValueError(f"precision must be one of {sorted(PRECISION)}, got {precision!r}")
    p = native.Params()
    p.nx, p.ny, p.nz = rp.nx, rp.ny, rp.nz
    p.jper, p.kper = rp.jper, rp.kper
    p.wall_x_min, p.wall_x_max = rp.wall_x
    p.wall_y_min, p.wall_y_max = rp.wall_y
    p.wall_z_min, p.wall_z_max = rp.wall_z
    p.inlet_bc, p.outlet_bc = rp.inlet_bc, rp.outlet_bc
    p.porous_plate_cmd, p.z_porous_plate = rp.porous_plate_cmd, rp.z_porous_plate
    p.n_exclude_inlet, p.n_exclude_outlet = rp.n_exclude_inlet, rp.n_exclude_outlet
    p.precision = PRECISION[precision]
    p.la_nui1 = 1.0 / rp.la_nu1
    p.la_nui2 = 1.0 / rp.la_nu2
    p.gamma, p.beta, p.force_z = rp.gamma, rp.beta, rp.force_z
    p.sa_inject, p.phi_inlet = rp.sa_inject, rp.phi_inlet
    p.rho_in, p.rho_out, p.uin_avg = rp.rho_in, rp.rho_out, rp.uin_avg
    p.relaxation = 1.0
    return p


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


class Case:
    """``geometry="device"`` builds the boundary lists on the GPU (mflbm_geometry_new_device, identical
    results); the default is the host routine.  ``boundary_file`` loads the lists from a boundary-info file
    instead (geometry_preprocess_cmd 1, Geometry_preprocessing.F90:524-685)."""

    def __init__(self, control: Dict[str, object], sample: np.ndarray, geometry: str = "host",
                 boundary_file: Optional[str] = None, precision: str = "fp64"):
        lib = native.load()
        self.control = dict(control)
        self.rp = io.derive_params(control)
        rp = self.rp
        nx, ny, nz = rp.nx, rp.ny, rp.nz
        self.walls_global = io.global_walls(sample, nx, ny, nz, rp.wall_x, rp.wall_y, rp.wall_z)
        self.precision = precision
        self.params = make_params(rp, precision)
        self.walls_h2 = np.zeros((nz + 4, ny + 4, nx + 4), dtype=np.int8)
        wg = np.ascontiguousarray(self.walls_global)
        _check(lib.mflbm_setup_walls(C.byref(self.params), i8ptr(wg), i8ptr(self.walls_h2)), "mflbm_setup_walls")
        if (rp.kper == 1 and nz < 10) or (rp.jper == 1 and ny < 10):
            # the geometry pass wraps 10 ghost layers around periodic directions
            # (Geometry_preprocessing.F90:54-83), which needs at least 10 nodes along them
            raise ValueError("a periodic direction needs at least 10 lattice nodes")
        if boundary_file is not None:
            s_ijk, s_nfl, s_nbr, s_w, f_ijk, f_nw = io.read_boundary_info(boundary_file, nx, ny, nz)
            b = BoundaryLists(s_ijk, s_nfl, s_nbr, s_w, f_ijk, f_nw, np.full(len(f_ijk), rp.theta))
        else:
            b = self._compute_lists(lib, wg, geometry)
        self.lists = b
        self.w_in = np.zeros((ny + 2, nx + 2))
        if rp.open_z and rp.inlet_bc == 1:
            _check(lib.mflbm_inlet_profile(C.byref(self.params), rp.uin_avg, rp.la_x, rp.la_y, dptr(self.w_in)),
                   "mflbm_inlet_profile")

    def _compute_lists(self, lib, wg: np.ndarray, geometry: str) -> "BoundaryLists":
        if geometry not in ("host", "device"):
            raise ValueError("geometry must be 'host' or 'device'")
        fn = lib.mflbm_geometry_new_device if geometry == "device" else lib.mflbm_geometry_new
        geo = C.c_void_p()
        _check(fn(C.byref(self.params), i8ptr(wg), self.rp.theta, C.byref(geo)), fn.__name__)
        try:
            ns, nf = C.c_int32(), C.c_int32()
            lib.mflbm_geometry_counts(geo, C.byref(ns), C.byref(nf))
            ns, nf = ns.value, nf.value
            b = BoundaryLists(np.zeros((ns, 3), np.int32), np.zeros(ns, np.int32), np.zeros((ns, 18), np.int32),
                              np.zeros(ns), np.zeros((nf, 3), np.int32), np.zeros((nf, 3)), np.zeros(nf))
            _check(lib.mflbm_geometry_export(geo, iptr(b.s_ijk), iptr(b.s_nfl), iptr(b.s_nbr), dptr(b.s_w),
                                             iptr(b.f_ijk), dptr(b.f_nw), dptr(b.f_theta)), "mflbm_geometry_export")
        finally:
            lib.mflbm_geometry_free(geo)
        return b

    @property
    def n_fluid(self) -> int:
        """pore_sum: fluid nodes of the interior (Misc.F90:298-365), the MLUPS node count."""
        return int((self.walls_global <= 0).sum())

    def initial_state(self, velocity: Optional[tuple] = None) -> State:
        """Fields of a new run; ``velocity`` = (u, v, w) arrays over the population shape to start from
        the equilibrium at that velocity instead of rest."""
        lib = native.load()
        rp = self.rp
        st = State.empty(rp.nx, rp.ny, rp.nz)
        fp = native.ptr_array(list(st.f))
        gp = native.ptr_array(list(st.g))
        uvw = [None, None, None]
        keep = []
        if velocity is not None:
            for m in range(3):
                a = np.ascontiguousarray(velocity[m], dtype=np.float64)
                keep.append(a)
                uvw[m] = dptr(a)
        _check(lib.mflbm_init_fields(C.byref(self.params), rp.init_option, rp.interface_z0, uvw[0], uvw[1], uvw[2],
                                     fp, gp, dptr(st.phi), dptr(st.f_convec), dptr(st.g_convec), dptr(st.phi_convec)),
               "mflbm_init_fields")
        return st
