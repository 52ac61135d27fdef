"""On-disk formats of the MF-LBM two-phase and single-phase solvers: control file, wall file, checkpoint file.

Formats kept byte-compatible with the reference so existing run directories keep working:

* control file ``simulation_control.txt``: one ``key value[,value...]`` pair per line, ``#`` comments;
  the parser takes the first whitespace-delimited token as the key and reads the rest list-directed
  (multiphase_3D/0.src/IO_multiphase.F90:73-278).
* wall file: stream of ``int32 nx, ny, nz`` then ``int8 walls(nx,ny,nz)`` x-fastest
  (multiphase_3D/0.src/Misc.F90:264-271, preprocessing/1.create_geometry_to_WallArray/sample_code_3d_geometry.f90:75-77).
* checkpoint ``out2.checkpoint/id0000``: stream of ``int32 ntime+1, f64 force_z, f64 rho_in``, then
  ``f0..f18, g0..g18`` each over ``(0:nx+1, 0:ny+1, 0:nz+1)``, ``phi`` over ``(-3:nx+4, -3:ny+4, -3:nz+4)``
  and, with the convective outlet, ``f_convec_bc, g_convec_bc`` over ``(0:nx+1, 0:ny+1, 0:18)`` and
  ``phi_convec_bc`` over ``(0:nx+1, 0:ny+1)`` (IO_multiphase.F90:586-631, read back at
  Init_multiphase.F90:499-545).

Arrays are returned as numpy arrays indexed ``[k, j, i]`` (C order, x fastest), so their memory is
exactly the reference's column-major layout.
"""
from __future__ import annotations

import dataclasses
import math
import os
import struct
from typing import Dict, Optional

import numpy as np

PI = 3.14159265358979323846

# -------------------------------------------------------------------------------------------------
# control file
# -------------------------------------------------------------------------------------------------

_INT_KEYS = {
    "initial_fluid_distribution_option": 1, "breakthrough_check": 1, "steady_state_option": 1,
    "benchmark_cmd": 1, "output_fieldData_precision_cmd": 1, "extreme_large_sim_cmd": 1,
    "modify_geometry_cmd": 1, "external_geometry_read_cmd": 1, "geometry_preprocess_cmd": 1,
    "porous_plate_cmd": 1, "change_inlet_fluid_phase_cmd": 1, "lattice_dimensions": 3,
    "excluded_layers": 2, "domain_wall_status_x": 2, "domain_wall_status_y": 2,
    "domain_wall_status_z": 2, "periodic_indicator": 3, "MPI_async_layers_num": 3,
    "MPI_process_num": 3, "inlet_BC": 1, "outlet_BC": 1, "Z_porous_plate": 1, "max_time_step": 1,
    "max_time_step_benchmark": 1, "ntime_visual": 1, "ntime_animation": 1, "monitor_timer": 1,
    "monitor_profile_timer_ratio": 1, "computation_time_timer": 1, "display_steps_timer": 1,
}
_REAL_KEYS = {
    "convergence_criteria", "fluid1_viscosity", "fluid2_viscosity", "surface_tension", "theta",
    "RK_beta", "target_inject_pore_volume", "capillary_number", "saturation_injection",
    "initial_interface_position", "body_force_0", "body_force_increment", "target_fluid1_saturation",
    "checkpoint_save_timer", "checkpoint_2rd_save_timer", "simulation_duration_timer",
    "d_vol_animation", "d_vol_detail", "d_vol_monitor",
}

DEFAULT_CONTROL: Dict[str, object] = {
    "initial_fluid_distribution_option": 1, "benchmark_cmd": 0, "extreme_large_sim_cmd": 0,
    "breakthrough_check": 0, "steady_state_option": 0, "convergence_criteria": 1e-6,
    "output_fieldData_precision_cmd": 1, "modify_geometry_cmd": 0, "external_geometry_read_cmd": 1,
    "geometry_preprocess_cmd": 0, "porous_plate_cmd": 0, "Z_porous_plate": 0,
    "change_inlet_fluid_phase_cmd": 0, "lattice_dimensions": (16, 16, 32), "excluded_layers": (0, 0),
    "domain_wall_status_x": (1, 1), "domain_wall_status_y": (1, 1), "domain_wall_status_z": (0, 0),
    "periodic_indicator": (0, 0, 1), "MPI_process_num": (1, 1, 1), "MPI_async_layers_num": (0, 0, 4),
    "fluid1_viscosity": 1.0 / 15.0, "fluid2_viscosity": 1.0 / 15.0, "surface_tension": 0.01,
    "theta": 90.0, "RK_beta": 0.95, "inlet_BC": 1, "outlet_BC": 1, "saturation_injection": 1.0,
    "target_inject_pore_volume": -1.0, "initial_interface_position": 20.0, "capillary_number": 0.0,
    "body_force_0": 1e-6, "body_force_increment": 0.0, "target_fluid1_saturation": 0.4,
    "max_time_step": 100, "max_time_step_benchmark": 100, "ntime_visual": 100000000,
    "ntime_animation": 100000000, "monitor_timer": 100000000, "monitor_profile_timer_ratio": 1000,
    "computation_time_timer": 100000000, "display_steps_timer": 100000000,
    "checkpoint_save_timer": 1000.0, "checkpoint_2rd_save_timer": 1000.0,
    "simulation_duration_timer": 1000.0, "d_vol_animation": -1.0, "d_vol_detail": -1.0,
    "d_vol_monitor": -1.0,
}


def _fmt(v) -> str:
    if isinstance(v, (tuple, list)):
        return ",".join(_fmt(x) for x in v)
    if isinstance(v, float):
        return repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
    return str(v)


def write_control(path: str, values: Dict[str, object]) -> None:
    """Write a control file with every key the reference parser knows (defaults overridden by values)."""
    merged = dict(DEFAULT_CONTROL)
    merged.update(values)
    with open(path, "w") as fh:
        fh.write("# control file written by mflbm.io.write_control\n")
        for k, v in merged.items():
            fh.write(f"{k} {_fmt(v)}\n")


def _parse_real(tok: str) -> float:
    return float(tok.strip().lower().replace("d", "e"))


def read_control(path: str) -> Dict[str, object]:
    """Parse a control file the way IO_multiphase.F90:73-278 does: first token is the key; unknown keys
    are ignored; values are list-directed (comma or blank separated, Fortran ``d`` exponents allowed)."""
    out: Dict[str, object] = {}
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            parts = line.split(None, 1)
            key = parts[0]
            rest = parts[1] if len(parts) > 1 else ""
            toks = [t for t in rest.replace(",", " ").split() if t]
            if key in _INT_KEYS:
                n = _INT_KEYS[key]
                vals = [int(t) for t in toks[:n]]
                out[key] = vals[0] if n == 1 else tuple(vals)
            elif key in _REAL_KEYS:
                out[key] = _parse_real(toks[0])
    return out


# -------------------------------------------------------------------------------------------------
# derived run parameters (initialization_basic_multi, Init_multiphase.F90:5-191 and IO_multiphase.F90:457-552)
# -------------------------------------------------------------------------------------------------

@dataclasses.dataclass
class RunParams:
    nx: int
    ny: int
    nz: int
    jper: int
    kper: int
    wall_x: tuple
    wall_y: tuple
    wall_z: tuple
    inlet_bc: int
    outlet_bc: int
    la_nu1: float
    la_nu2: float
    gamma: float
    beta: float
    theta: float            # radians, measured through fluid 2 (IO_multiphase.F90:467-468)
    force_z: float
    sa_inject: float
    phi_inlet: float
    rho_in: float
    rho_out: float
    uin_avg: float
    la_x: float
    la_y: float
    A_xy: float
    init_option: int
    interface_z0: float
    n_exclude_inlet: int
    n_exclude_outlet: int
    porous_plate_cmd: int = 0
    z_porous_plate: int = 0

    @property
    def open_z(self) -> bool:
        """Non-periodic z with open ends: the inlet/outlet boundary conditions are active."""
        return self.kper == 0 and self.wall_z[0] == 0 and self.wall_z[1] == 0


def _f32(x: float) -> float:
    return float(np.float32(x))


def derive_params(ctl: Dict[str, object]) -> RunParams:
    """Derive the per-run constants exactly as the reference does (single-rank case)."""
    c = dict(DEFAULT_CONTROL)
    c.update(ctl)
    nx, ny, nz = c["lattice_dimensions"]
    iper, jper, kper = c["periodic_indicator"]
    wx, wy, wz = tuple(c["domain_wall_status_x"]), tuple(c["domain_wall_status_y"]), tuple(c["domain_wall_status_z"])
    # validation, IO_multiphase.F90:471-552
    if iper == 1 or wx[0] == 0 or wx[1] == 0:
        raise ValueError("x must carry solid walls at both ends and cannot be periodic")
    if jper == 0 and (wy[0] == 0 or wy[1] == 0):
        raise ValueError("y without periodicity needs solid walls at both ends")
    if jper == 1 and (wy[0] == 1 or wy[1] == 1):
        raise ValueError("y periodic with solid y walls")
    if kper == 1 and (wz[0] == 1 or wz[1] == 1):
        raise ValueError("z periodic with solid z walls")
    inlet_bc, outlet_bc = int(c["inlet_BC"]), int(c["outlet_BC"])
    if outlet_bc == 1 and inlet_bc == 2:
        raise ValueError("pressure inlet with convective outlet is not supported")
    theta = 180.0 - float(c["theta"])
    theta = theta * PI / 180.0
    la_nu1, la_nu2 = float(c["fluid1_viscosity"]), float(c["fluid2_viscosity"])
    gamma = float(c["surface_tension"])
    sa_inject = float(c["saturation_injection"])
    phi_inlet = 2.0 * sa_inject - 1.0
    force_z0 = float(c["body_force_0"])
    force_z = force_z0
    rho_out = 1.0
    rho_in = 0.0
    uin_avg = 0.0
    # la_x, la_y are evaluated in default (single) precision in the reference: nxGlobal - 1 - 0.5 - 0.5
    la_x = _f32(_f32(_f32(nx - 1) - 0.5) - 0.5)
    la_y = _f32(_f32(_f32(ny - 1) - 0.5) - 0.5)
    A_xy = la_x * la_y
    open_z = kper == 0 and wz[0] == 0 and wz[1] == 0
    if open_z:
        if inlet_bc == 1:
            force_z = 0.0
            uin_avg = float(c["capillary_number"]) * gamma / la_nu1
        elif inlet_bc == 2:
            p_gradient = -force_z0 / 3.0
            rho_in = rho_out - p_gradient * nz
    return RunParams(
        nx=nx, ny=ny, nz=nz, jper=jper, kper=kper, wall_x=wx, wall_y=wy, wall_z=wz,
        inlet_bc=inlet_bc, outlet_bc=outlet_bc, la_nu1=la_nu1, la_nu2=la_nu2, gamma=gamma,
        beta=float(c["RK_beta"]), theta=theta, force_z=force_z, sa_inject=sa_inject,
        phi_inlet=phi_inlet, rho_in=rho_in, rho_out=rho_out, uin_avg=uin_avg, la_x=la_x, la_y=la_y,
        A_xy=A_xy, init_option=int(c["initial_fluid_distribution_option"]),
        interface_z0=float(c["initial_interface_position"]),
        n_exclude_inlet=int(c["excluded_layers"][0]), n_exclude_outlet=int(c["excluded_layers"][1]),
        porous_plate_cmd=int(c["porous_plate_cmd"]), z_porous_plate=int(c["Z_porous_plate"]),
    )


# -------------------------------------------------------------------------------------------------
# wall file
# -------------------------------------------------------------------------------------------------

def write_walls(path: str, walls: np.ndarray) -> None:
    """walls: int8 array indexed [k, j, i] of shape (nz, ny, nx); 1 = solid, 0 = fluid."""
    nz, ny, nx = walls.shape
    with open(path, "wb") as fh:
        fh.write(struct.pack("<3i", nx, ny, nz))
        fh.write(np.ascontiguousarray(walls, dtype=np.int8).tobytes())


def read_walls(path: str) -> np.ndarray:
    with open(path, "rb") as fh:
        nx, ny, nz = struct.unpack("<3i", fh.read(12))
        data = np.frombuffer(fh.read(nx * ny * nz), dtype=np.int8)
    return data.reshape(nz, ny, nx).copy()


def global_walls(sample: np.ndarray, nx: int, ny: int, nz: int, wall_x, wall_y, wall_z) -> np.ndarray:
    """walls_global as set_walls builds it for an external geometry (Misc.F90:18-109, read_walls :247-295):
    the sample is placed at the origin, padded with solid when both x_max and y_max walls are on, then
    the domain walls are stamped."""
    nzs, nys, nxs = sample.shape
    if nx < nxs or ny < nys or nz < nzs:
        raise ValueError("domain smaller than the geometry sample")
    g = np.zeros((nz, ny, nx), dtype=np.int8)
    g[:nzs, :nys, :nxs] = sample
    if wall_x[1] == 1 and wall_y[1] == 1:
        i = np.arange(1, nx + 1)[None, None, :]
        j = np.arange(1, ny + 1)[None, :, None]
        pad = (j >= nys) | (i >= nxs)
        g[np.broadcast_to(pad, g.shape)] = 1
    if wall_z[0] == 1:
        g[0, :, :] = 1
    if wall_z[1] == 1:
        g[-1, :, :] = 1
    if wall_x[0] == 1:
        g[:, :, 0] = 1
    if wall_x[1] == 1:
        g[:, :, -1] = 1
    if wall_y[0] == 1:
        g[:, 0, :] = 1
    if wall_y[1] == 1:
        g[:, -1, :] = 1
    return g


# -------------------------------------------------------------------------------------------------
# checkpoint
# -------------------------------------------------------------------------------------------------

@dataclasses.dataclass
class Checkpoint:
    ntime_next: int
    force_z: float
    rho_in: float
    f: np.ndarray              # (19, nz+2, ny+2, nx+2)
    g: np.ndarray              # (19, nz+2, ny+2, nx+2)
    phi: np.ndarray            # (nz+8, ny+8, nx+8)
    f_convec: Optional[np.ndarray] = None   # (19, ny+2, nx+2)
    g_convec: Optional[np.ndarray] = None
    phi_convec: Optional[np.ndarray] = None  # (ny+2, nx+2)


def read_checkpoint(path: str, nx: int, ny: int, nz: int, convective: bool) -> Checkpoint:
    npdf = (nx + 2) * (ny + 2) * (nz + 2)
    nphi = (nx + 8) * (ny + 8) * (nz + 8)
    nc = (nx + 2) * (ny + 2)
    with open(path, "rb") as fh:
        raw = fh.read()
    ntime_next, force_z, rho_in = struct.unpack("<idd", raw[:20])
    off = 20
    arr = np.frombuffer(raw, dtype="<f8", offset=off, count=38 * npdf).reshape(38, nz + 2, ny + 2, nx + 2)
    off += 38 * npdf * 8
    phi = np.frombuffer(raw, dtype="<f8", offset=off, count=nphi).reshape(nz + 8, ny + 8, nx + 8)
    off += nphi * 8
    ck = Checkpoint(ntime_next, force_z, rho_in, arr[:19].copy(), arr[19:].copy(), phi.copy())
    if convective:
        cf = np.frombuffer(raw, dtype="<f8", offset=off, count=19 * nc).reshape(19, ny + 2, nx + 2)
        off += 19 * nc * 8
        cg = np.frombuffer(raw, dtype="<f8", offset=off, count=19 * nc).reshape(19, ny + 2, nx + 2)
        off += 19 * nc * 8
        cp = np.frombuffer(raw, dtype="<f8", offset=off, count=nc).reshape(ny + 2, nx + 2)
        off += nc * 8
        ck.f_convec, ck.g_convec, ck.phi_convec = cf.copy(), cg.copy(), cp.copy()
    if off != len(raw):
        raise ValueError(f"checkpoint size mismatch: parsed {off} of {len(raw)} bytes")
    return ck


def write_checkpoint(path: str, ck: Checkpoint) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<idd", ck.ntime_next, ck.force_z, ck.rho_in))
        fh.write(np.ascontiguousarray(ck.f, dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(ck.g, dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(ck.phi, dtype="<f8").tobytes())
        if ck.f_convec is not None:
            fh.write(np.ascontiguousarray(ck.f_convec, dtype="<f8").tobytes())
            fh.write(np.ascontiguousarray(ck.g_convec, dtype="<f8").tobytes())
            fh.write(np.ascontiguousarray(ck.phi_convec, dtype="<f8").tobytes())


# -------------------------------------------------------------------------------------------------
# text output in Fortran edit descriptors (the monitor files, Monitor.F90:153-248; info.txt, time.dat)
# -------------------------------------------------------------------------------------------------

def fortran_e(x: float, w: int, d: int) -> str:
    """Ew.d edit of x: ``0.<d digits>E<sign><2 digits>`` right-justified in w characters (three-digit
    exponents drop the ``E``, as the Fortran standard prescribes; NaN and infinities as the runtime prints
    them)."""
    if math.isnan(x):
        return f"{'NaN':>{w}s}"
    if math.isinf(x):
        return f"{('-' if x < 0 else '') + 'Infinity':>{w}s}"
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
        return f"{('-' if math.copysign(1.0, x) < 0 else '') + body:>{w}s}"
    m, e = f"{abs(x):.{d - 1}e}".split("e")
    digits = m.replace(".", "")
    ex = int(e) + 1
    exp = f"E{'+' if ex >= 0 else '-'}{abs(ex):02d}" if abs(ex) <= 99 else f"{'+' if ex >= 0 else '-'}{abs(ex):03d}"
    s = ("-" if x < 0 else "") + "0." + digits + exp
    return f"{s:>{w}s}" if len(s) <= w else "*" * w


def fortran_f(x: float, w: int, d: int) -> str:
    """Fw.d edit of x right-justified in w characters (asterisks when it does not fit)."""
    if math.isnan(x):
        return f"{'NaN':>{w}s}"
    s = f"{x:.{d}f}"
    if len(s) > w and s.startswith("0."):
        s = s[1:]          # the optional leading zero is dropped first
    elif len(s) > w and s.startswith("-0."):
        s = "-" + s[2:]
    return f"{s:>{w}s}" if len(s) <= w else "*" * w


_EDIT = {"e14.7": (14, 7), "E14.6": (14, 6)}


def fortran_fields(edit: str, *values: float) -> str:
    """``n(1x, Ew.d)`` of the values."""
    w, d = _EDIT[edit]
    return "".join(" " + fortran_e(float(v), w, d) for v in values)


def fortran_line(ntime: int, edit: str, *values: float) -> str:
    """``(I10, n(1x, Ew.d))``: the step and the values, one monitor-file line."""
    return f"{ntime:10d}" + fortran_fields(edit, *values)


def profile_name(ntime: int) -> str:
    """File name of a monitor profile: ``write(flnm,'(I10.8,".dat")')`` (Monitor.F90:229-231), i.e. the step
    with at least 8 digits right-justified in 10 characters, then ``.dat``."""
    return f"{ntime:10d}.dat" if ntime >= 10 ** 8 else "  " + f"{ntime:08d}.dat"


# =================================================================================================
# single-phase program (singlephase_3D/0.src): control file, derived constants, checkpoint
# =================================================================================================

_SP_INT_KEYS = {
    "steady_state_option": 1, "MRT_collision_parameter_preset": 1, "benchmark_cmd": 1,
    "output_fieldData_precision_cmd": 1, "extreme_large_sim_cmd": 1, "modify_geometry_cmd": 1,
    "external_geometry_read_cmd": 1, "lattice_dimensions": 3, "excluded_layers": 2,
    "domain_wall_status_x": 2, "domain_wall_status_y": 2, "domain_wall_status_z": 2,
    "periodic_indicator": 3, "MPI_async_layers_num": 3, "MPI_process_num": 3, "inlet_BC": 1,
    "outlet_BC": 1, "max_time_step": 1, "max_time_step_benchmark": 1, "ntime_visual": 1,
    "monitor_timer": 1, "monitor_profile_timer_ratio": 1, "computation_time_timer": 1,
    "display_steps_timer": 1,
}
_SP_REAL_KEYS = {
    "convergence_criteria", "char_length", "fluid_viscosity", "body_force_0", "body_force_increment",
    "Reynolds_number", "rho_drop", "target_inject_pore_volume", "checkpoint_save_timer",
    "checkpoint_2rd_save_timer", "simulation_duration_timer", "d_vol_detail", "d_vol_monitor",
}

SP_DEFAULT_CONTROL: Dict[str, object] = {
    "benchmark_cmd": 0, "extreme_large_sim_cmd": 0, "steady_state_option": 0, "MRT_collision_parameter_preset": 1,
    "convergence_criteria": 1e-6, "modify_geometry_cmd": 0, "external_geometry_read_cmd": 1,
    "output_fieldData_precision_cmd": 1, "lattice_dimensions": (16, 16, 32), "excluded_layers": (0, 0),
    "domain_wall_status_x": (1, 1), "domain_wall_status_y": (1, 1), "domain_wall_status_z": (0, 0),
    "periodic_indicator": (0, 0, 1), "MPI_process_num": (1, 1, 1), "MPI_async_layers_num": (0, 0, 4),
    "char_length": 20.0, "Reynolds_number": 10.0, "rho_drop": 0.01, "fluid_viscosity": 0.05, "inlet_BC": 0,
    "outlet_BC": 0, "target_inject_pore_volume": -1.0, "body_force_0": 1e-6, "body_force_increment": 0.0,
    "max_time_step": 100, "max_time_step_benchmark": 100, "ntime_visual": 100000000, "monitor_timer": 100000000,
    "monitor_profile_timer_ratio": 1000, "computation_time_timer": 100000000, "display_steps_timer": 100000000,
    "checkpoint_save_timer": 1000.0, "checkpoint_2rd_save_timer": 1000.0, "simulation_duration_timer": 1000.0,
    "d_vol_detail": -1.0, "d_vol_monitor": -1.0,
}


def write_control_sp(path: str, values: Dict[str, object]) -> None:
    """Write a single-phase control file with every key its parser knows (IO.F90:65-290)."""
    merged = dict(SP_DEFAULT_CONTROL)
    merged.update(values)
    with open(path, "w") as fh:
        fh.write("# control file written by mflbm.io.write_control_sp\n")
        for k, v in merged.items():
            fh.write(f"{k} {_fmt(v)}\n")


def read_control_sp(path: str) -> Dict[str, object]:
    """Parse a single-phase control file (IO.F90:65-290); same line grammar as the two-phase parser."""
    out: Dict[str, object] = {}
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            parts = line.split(None, 1)
            key = parts[0]
            toks = [t for t in (parts[1] if len(parts) > 1 else "").replace(",", " ").split() if t]
            if key in _SP_INT_KEYS:
                n = _SP_INT_KEYS[key]
                vals = [int(t) for t in toks[:n]]
                out[key] = vals[0] if n == 1 else tuple(vals)
            elif key in _SP_REAL_KEYS:
                out[key] = _parse_real(toks[0])
    return out


@dataclasses.dataclass
class SPRunParams:
    nx: int
    ny: int
    nz: int
    jper: int
    kper: int
    wall_x: tuple
    wall_y: tuple
    wall_z: tuple
    inlet_bc: int
    outlet_bc: int
    la_nu: float
    mrt_preset: int
    force_z: float
    rho_in: float
    rho_out: float
    uin_avg: float
    flowrate0: float        # inlet flow rate of a velocity inlet (uin_avg * A_xy)
    la_x: float
    la_y: float
    A_xy: float
    n_exclude_inlet: int
    n_exclude_outlet: int
    steady_state_option: int
    convergence_criteria: float

    @property
    def open_z(self) -> bool:
        return self.kper == 0 and self.wall_z[0] == 0 and self.wall_z[1] == 0


def derive_sp_params(ctl: Dict[str, object]) -> SPRunParams:
    """Per-run constants of the single-phase program (read_parameter, IO.F90:318-484, and
    initialization_basic / initialization_open_velocity_inlet_BC, Initialization.F90:4-205)."""
    c = dict(SP_DEFAULT_CONTROL)
    c.update(ctl)
    nx, ny, nz = c["lattice_dimensions"]
    iper, jper, kper = c["periodic_indicator"]
    wx, wy, wz = tuple(c["domain_wall_status_x"]), tuple(c["domain_wall_status_y"]), tuple(c["domain_wall_status_z"])
    if iper == 1 or wx[0] == 0 or wx[1] == 0:
        raise ValueError("x must carry solid walls at both ends and cannot be periodic")
    if jper == 0 and (wy[0] == 0 or wy[1] == 0):
        raise ValueError("y without periodicity needs solid walls at both ends")
    if jper == 1 and (wy[0] == 1 or wy[1] == 1):
        raise ValueError("y periodic with solid y walls")
    if kper == 1 and (wz[0] == 1 or wz[1] == 1):
        raise ValueError("z periodic with solid z walls")
    inlet_bc, outlet_bc = int(c["inlet_BC"]), int(c["outlet_BC"])
    if outlet_bc == 1 and inlet_bc == 2:
        raise ValueError("pressure inlet with convective outlet is not supported")
    la_nu = float(c["fluid_viscosity"])
    la_x = _f32(_f32(_f32(nx - 1) - 0.5) - 0.5)
    la_y = _f32(_f32(_f32(ny - 1) - 0.5) - 0.5)
    A_xy = la_x * la_y
    rho_out, rho_in, uin_avg, flowrate0 = 1.0, 0.0, 0.0, 0.0
    open_z = kper == 0 and wz[0] == 0 and wz[1] == 0
    if open_z:
        if inlet_bc == 1:
            uin_avg = float(c["Reynolds_number"]) * la_nu / float(c["char_length"])
            flowrate0 = uin_avg * A_xy
        elif inlet_bc == 2:
            rho_in = rho_out + float(c["rho_drop"])
    return SPRunParams(
        nx=nx, ny=ny, nz=nz, jper=jper, kper=kper, wall_x=wx, wall_y=wy, wall_z=wz, inlet_bc=inlet_bc,
        outlet_bc=outlet_bc, la_nu=la_nu, mrt_preset=int(c["MRT_collision_parameter_preset"]),
        force_z=float(c["body_force_0"]), rho_in=rho_in, rho_out=rho_out, uin_avg=uin_avg, flowrate0=flowrate0,
        la_x=la_x, la_y=la_y, A_xy=A_xy, n_exclude_inlet=int(c["excluded_layers"][0]),
        n_exclude_outlet=int(c["excluded_layers"][1]), steady_state_option=int(c["steady_state_option"]),
        convergence_criteria=float(c["convergence_criteria"]))


@dataclasses.dataclass
class SPCheckpoint:
    ntime_next: int
    force_z: float
    rho_in: float
    f: np.ndarray                           # (19, nz+2, ny+2, nx+2)
    f_convec: Optional[np.ndarray] = None   # (19, ny+2, nx+2)


def read_checkpoint_sp(path: str, nx: int, ny: int, nz: int, convective: bool) -> SPCheckpoint:
    """Single-phase checkpoint (save_checkpoint, IO.F90:490-547): int32 ntime+1, f64 force_z, f64 rho_in,
    f0..f18 over (0:nx+1, 0:ny+1, 0:nz+1), then f_convec_bc over (0:nx+1, 0:ny+1, 0:18) when outlet_BC = 1."""
    npdf = (nx + 2) * (ny + 2) * (nz + 2)
    nc = (nx + 2) * (ny + 2)
    with open(path, "rb") as fh:
        raw = fh.read()
    ntime_next, force_z, rho_in = struct.unpack("<idd", raw[:20])
    off = 20
    f = np.frombuffer(raw, dtype="<f8", offset=off, count=19 * npdf).reshape(19, nz + 2, ny + 2, nx + 2)
    off += 19 * npdf * 8
    ck = SPCheckpoint(ntime_next, force_z, rho_in, f.copy())
    if convective:
        cf = np.frombuffer(raw, dtype="<f8", offset=off, count=19 * nc).reshape(19, ny + 2, nx + 2)
        off += 19 * nc * 8
        ck.f_convec = cf.copy()
    if off != len(raw):
        raise ValueError(f"checkpoint size mismatch: parsed {off} of {len(raw)} bytes")
    return ck


def write_checkpoint_sp(path: str, ck: SPCheckpoint) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<idd", ck.ntime_next, ck.force_z, ck.rho_in))
        fh.write(np.ascontiguousarray(ck.f, dtype="<f8").tobytes())
        if ck.f_convec is not None:
            fh.write(np.ascontiguousarray(ck.f_convec, dtype="<f8").tobytes())


# -------------------------------------------------------------------------------------------------
# boundary-info file (geometry_preprocess_cmd 1): the lists written by the reference's preprocessing tool
# (preprocessing/3.wall_boundary_preprocess/3d_wall_boundary_process.f90) and read by
# geometry_preprocessing_load (multiphase_3D/0.src/Geometry_preprocessing.F90:524-685)
# -------------------------------------------------------------------------------------------------

# stream layout: int32 n_solid, n_fluid; n_solid records {int32 ix, iy, iz, i_fluid_num, neighbor_list(18);
# f64 la_weight} (96 bytes); n_fluid records {int32 ix, iy, iz; f64 nwx, nwy, nwz}, 40 bytes where the
# compiler aligns the reals (gfortran: 4 bytes of padding) or 36 bytes packed (flang).  The record layout
# follows the compiler of the tool that wrote the file (its compile.sh says so); both are read.
_SOLID_REC = np.dtype([("ijk", "<i4", 3), ("nfl", "<i4"), ("nbr", "<i4", 18), ("w", "<f8")])
_FLUID_REC = np.dtype([("ijk", "<i4", 3), ("pad", "<i4"), ("nw", "<f8", 3)])
_FLUID_REC_PACKED = np.dtype([("ijk", "<i4", 3), ("nw", "<f8", 3)], align=False)
assert _SOLID_REC.itemsize == 96 and _FLUID_REC.itemsize == 40 and _FLUID_REC_PACKED.itemsize == 36


def write_boundary_info(path: str, s_ijk, s_nfl, s_nbr, s_w, f_ijk, f_nw, packed: bool = False) -> None:
    s = np.zeros(len(s_nfl), dtype=_SOLID_REC)
    s["ijk"], s["nfl"], s["nbr"], s["w"] = s_ijk, s_nfl, s_nbr, s_w
    f = np.zeros(len(f_ijk), dtype=_FLUID_REC_PACKED if packed else _FLUID_REC)
    f["ijk"], f["nw"] = f_ijk, f_nw
    with open(path, "wb") as fh:
        fh.write(struct.pack("<ii", len(s), len(f)))
        fh.write(s.tobytes())
        fh.write(f.tobytes())


def read_boundary_info(path: str, nx: int, ny: int, nz: int):
    """Read a boundary-info file and keep, as geometry_preprocessing_load does for a single rank, the solid
    nodes within 3 ghost layers and the fluid nodes within 2 (:606-676).  Returns (s_ijk, s_nfl, s_nbr, s_w,
    f_ijk, f_nw) as contiguous arrays; the caller assigns the contact angle to every fluid node."""
    with open(path, "rb") as fh:
        raw = fh.read()
    ns, nf = struct.unpack("<ii", raw[:8])
    rest = len(raw) - 8 - ns * 96
    if rest == nf * 40:
        frec = _FLUID_REC
    elif rest == nf * 36:
        frec = _FLUID_REC_PACKED
    else:
        raise ValueError(f"boundary-info file size does not match {ns} solid and {nf} fluid records")
    s = np.frombuffer(raw, dtype=_SOLID_REC, count=ns, offset=8)
    f = np.frombuffer(raw, dtype=frec, count=nf, offset=8 + ns * 96)

    def keep(ijk, g):
        lo = 1 - g
        return ((ijk[:, 0] >= lo) & (ijk[:, 0] <= nx + g) & (ijk[:, 1] >= lo) & (ijk[:, 1] <= ny + g) &
                (ijk[:, 2] >= lo) & (ijk[:, 2] <= nz + g))

    ks, kf = keep(s["ijk"], 3), keep(f["ijk"], 2)
    c = np.ascontiguousarray
    return (c(s["ijk"][ks], dtype=np.int32), c(s["nfl"][ks], dtype=np.int32), c(s["nbr"][ks], dtype=np.int32),
            c(s["w"][ks], dtype=np.float64), c(f["ijk"][kf], dtype=np.int32), c(f["nw"][kf], dtype=np.float64))
