"""Single-phase runs on the GPU (singlephase_3D/0.src of the reference).

* ``SPCase(control, sample)``: per-run constants (read_parameter / initialization_basic,
  IO.F90:318-484, Initialization.F90:4-205), local walls with halo 2 (set_walls), the inlet profile and the
  initial populations (initialization_new, Initialization.F90:207-307).
* ``SPSimulation``: the device state behind the ``mflbm_sp_*`` C ABI: step / run, monitor (Monitor.F90:4-172),
  checkpoint state, and ``run_steady``, the reference's main loop with its flow-rate convergence test
  (Main.F90:124-230, Monitor.F90:145-168).
* ``permeability``: k = nu * Q / (A_xy * F) for a body-force-driven sample (the reference stops at the flow
  rate; the permeability is the quantity its users derive from it).
* ``duct_flow_rate``: the analytic volumetric flow rate of a rectangular duct under a body force, the series
  behind inlet_vel_profile_rectangular (Misc.F90:493-532), used to check the solver physics.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

from . import io, native
from .native import dptr, i8ptr

EPS = 1e-14   # eps of the single-phase program (Module.F90:6)


@dataclasses.dataclass
class SPState:
    f: np.ndarray          # (19, nz+2, ny+2, nx+2)
    f_convec: np.ndarray   # (19, ny+2, nx+2)

    @staticmethod
    def empty(nx: int, ny: int, nz: int) -> "SPState":
        return SPState(np.zeros((19, nz + 2, ny + 2, nx + 2)), np.zeros((19, ny + 2, nx + 2)))

    @staticmethod
    def from_checkpoint(ck: io.SPCheckpoint) -> "SPState":
        nz2, ny2, nx2 = ck.f.shape[1:]
        fc = ck.f_convec if ck.f_convec is not None else np.zeros((19, ny2, nx2))
        return SPState(np.ascontiguousarray(ck.f), np.ascontiguousarray(fc))


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


PRECISION = {"fp64": 0, "fp32": 1, "mixed": 2}

# lattice weights of the reference's equilibrium (singlephase_3D/0.src/Module.F90:93-98) and the rest state of
# its MRT collision: density 1, energy moment -11, energy-square moment 0, every other moment 0, mapped back to
# populations by the transform coefficients c1 = 1/19, c2 = 1/2394, c3 = 1/252 (Module.F90:116-125) with the
# same double operations the device evaluates, so the values are bit-identical to the device's (4/21, 13/126,
# 1/63 to rounding)
LATTICE_W = (1.0 / 3.0,) + (1.0 / 18.0,) * 6 + (1.0 / 36.0,) * 12
_C1, _C2, _C3, _E2C1 = 1.0 / 19.0, 1.0 / 2394.0, 1.0 / 252.0, 0.0
REST = ((_C1 + 330.0 * _C2 + 12.0 * _C3 * _E2C1,) + (_C1 + 121.0 * _C2 - 4.0 * _C3 * _E2C1,) * 6 +
        (_C1 - 88.0 * _C2 + _C3 * _E2C1,) * 12)


_EX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0)
_EY = (0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1)
_EZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1)
_OPP = (0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15)


def _bounce_back_solid_slots(f: np.ndarray, walls_global: np.ndarray) -> None:
    """Solid slot (y, q) facing a fluid node x = y + e_q takes x's population opp(q), in place.

    A fluid node pulls its population q from that slot at the first (odd) step; under bounce-back it is the
    population x sent toward the wall, f_opp(q)(x).  A started flow whose wall slots hold the rest state instead
    gets a momentum defect of the order of the flow itself at every wall link, and the pressure waves that
    defect launches decay as slowly as a start from rest.  f: (19, nz+2, ny+2, nx+2), interior at [1:-1]."""
    fl = np.zeros(f.shape[1:], dtype=bool)
    fl[1:-1, 1:-1, 1:-1] = walls_global == 0
    nz, ny, nx = fl.shape
    for q in range(1, 19):
        ez, ey, ex = _EZ[q], _EY[q], _EX[q]
        # y ranges over the nodes whose neighbour y + e_q lies inside the array
        ys = tuple(slice(max(0, -e), n - max(0, e)) for e, n in ((ez, nz), (ey, ny), (ex, nx)))
        xs = tuple(slice(max(0, e), n - max(0, -e)) for e, n in ((ez, nz), (ey, ny), (ex, nx)))
        m = ~fl[ys] & fl[xs]
        f[q][ys][m] = f[_OPP[q]][xs][m]


def make_sp_params(rp: io.SPRunParams, precision: str = "fp64") -> native.SPParams:
    """precision "fp64": the reference arithmetic; "fp32": populations stored in FP32 as deviations from the
    rest equilibrium and collided in FP32; "mixed": the same FP32 storage collided in FP64 (boundaries,
    macroscopic fields and reductions stay FP64 in every mode)."""
    if precision not in PRECISION:
        raise ValueError(f"precision must be one of {sorted(PRECISION)}, got {precision!r}")
    p = native.SPParams()
    p.nx, p.ny, p.nz = rp.nx, rp.ny, rp.nz
    p.jper, p.kper = rp.jper, rp.kper
    p.wall_x_min, p.wall_x_max = rp.wall_x
    p.wall_y_min, p.wall_y_max = rp.wall_y
    p.wall_z_min, p.wall_z_max = rp.wall_z
    p.inlet_bc, p.outlet_bc = rp.inlet_bc, rp.outlet_bc
    p.n_exclude_inlet, p.n_exclude_outlet = rp.n_exclude_inlet, rp.n_exclude_outlet
    p.mrt_preset = rp.mrt_preset
    p.precision = PRECISION[precision]
    p.la_nu, p.force_z = rp.la_nu, rp.force_z
    p.rho_in, p.rho_out, p.uin_avg = rp.rho_in, rp.rho_out, rp.uin_avg
    p.relaxation = 1.0
    return p


class SPCase:
    def __init__(self, control: Dict[str, object], sample: np.ndarray, precision: str = "fp64"):
        lib = native.load()
        self.control = dict(control)
        self.rp = io.derive_sp_params(control)
        self.precision = precision
        rp = self.rp
        nx, ny, nz = rp.nx, rp.ny, rp.nz
        self.walls_global = io.global_walls(sample, nx, ny, nz, rp.wall_x, rp.wall_y, rp.wall_z)
        self.params = make_sp_params(rp, precision)
        # set_walls and the inlet profile are the same routines in both programs; they read only the
        # lattice and wall fields of mflbm_params
        wp = native.Params()
        wp.nx, wp.ny, wp.nz, wp.jper, wp.kper = nx, ny, nz, rp.jper, rp.kper
        wp.wall_x_min, wp.wall_x_max = rp.wall_x
        wp.wall_y_min, wp.wall_y_max = rp.wall_y
        wp.wall_z_min, wp.wall_z_max = rp.wall_z
        self.walls_h2 = np.zeros((nz + 4, ny + 4, nx + 4), dtype=np.int8)
        wg = np.ascontiguousarray(self.walls_global)
        _check(lib.mflbm_setup_walls(C.byref(wp), i8ptr(wg), i8ptr(self.walls_h2)), "mflbm_setup_walls")
        self.w_in = np.zeros((ny + 2, nx + 2))
        if rp.open_z and rp.inlet_bc == 1:
            _check(lib.mflbm_inlet_profile(C.byref(wp), rp.uin_avg, rp.la_x, rp.la_y, dptr(self.w_in)),
                   "mflbm_inlet_profile")

    @property
    def pore_profile_z(self) -> np.ndarray:
        """Fluid nodes per z layer (pore_profile, Misc.F90:297-363)."""
        return (self.walls_global <= 0).sum(axis=(1, 2))

    @property
    def n_fluid(self) -> int:
        """pore_sum, the MLUPS node count."""
        return int((self.walls_global <= 0).sum())

    def initial_state(self, rho: Optional[np.ndarray] = None, velocity: Optional[tuple] = None) -> SPState:
        lib = native.load()
        rp = self.rp
        st = SPState.empty(rp.nx, rp.ny, rp.nz)
        keep = []
        args = [None, None, None, None]
        for m, a in enumerate([rho] + list(velocity if velocity is not None else (None, None, None))):
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float64)
                keep.append(a)
                args[m] = dptr(a)
        fc = st.f_convec if rp.outlet_bc == 1 else None
        _check(lib.mflbm_sp_init_fields(C.byref(self.params), args[0], args[1], args[2], args[3],
                                        native.ptr_array(list(st.f)), dptr(fc) if fc is not None else None),
               "mflbm_sp_init_fields")
        return st

    def collision_equilibrium_state(self, rho: Optional[np.ndarray] = None,
                                    velocity: Optional[tuple] = None) -> SPState:
        """Initial populations at the MRT collision's own equilibrium instead of the lattice-weight one.

        The reference's new-run state (``initial_state``, Initialization.F90:244-276) is the second-order
        equilibrium with the lattice weights (1/3, 1/18, 1/36), whose energy-square moment is 3 rho, while the
        collision relaxes that moment to 0 (e2 equilibrium without a density term): the rest state of the
        collision is REST = (4/21, 13/126, 1/63).  The first steps of a reference run therefore move mass
        between directions; at solid links this launches density waves of order 0.1, and their decay, not the
        flow's development, sets the time a run needs to meet the flow-rate criterion.  This state adds
        (REST_q - w_q) rho to every population (as (f_q - w_q rho) + REST_q rho, which is REST_q exactly at
        rest), so the rest state is a fixed point of the collision and a prescribed (rho, u) starts without
        that transient."""
        st = self.initial_state(rho=rho, velocity=velocity)
        r = np.ones_like(st.f[0]) if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        for q in range(19):
            st.f[q] = (st.f[q] - LATTICE_W[q] * r) + REST[q] * r
        if velocity is not None or rho is not None:
            _bounce_back_solid_slots(st.f, self.walls_global)
        return st


@dataclasses.dataclass
class SPMonitorResult:
    """Rank-0 quantities of monitor (Monitor.F90:84-168)."""
    flowrate: float                 # fl_avg_whole: mean over z of the layer flux sum(w)
    umax: float
    pressure: Optional[tuple]       # (mean rho at the first / last analysed layer, difference), open z only
    failed: bool
    layers: np.ndarray              # (2, nz): fl(k), pre(k)


def sp_post_monitor(case: SPCase, layers: np.ndarray, umax2: float) -> SPMonitorResult:
    rp = case.rp
    nz = rp.nz
    fl, pre = layers
    tot = 0.0
    for k in range(nz):   # serial, in layer order (Monitor.F90:110-116)
        tot += fl[k]
    flowrate = tot / nz
    umax = math.sqrt(umax2)
    pr = None
    if rp.open_z:
        pore = case.pore_profile_z
        k0, k1 = rp.n_exclude_inlet, nz - 1 - rp.n_exclude_outlet
        t1 = pre[k0] / pore[k0] if pore[k0] else float("nan")
        t2 = pre[k1] / pore[k1] if pore[k1] else float("nan")
        pr = (t1, t2, t1 - t2)
    failed = math.isnan(flowrate) or umax > 0.5
    return SPMonitorResult(flowrate, umax, pr, failed, layers)


def permeability(case: SPCase, flowrate: float) -> float:
    """Darcy permeability (lattice units) of a body-force-driven run: k = nu * Q / (A_xy * F)."""
    rp = case.rp
    if rp.force_z == 0.0:
        raise ValueError("permeability from the flow rate needs a body force")
    return rp.la_nu * flowrate / (rp.A_xy * rp.force_z)


def duct_flow_rate(force: float, nu: float, width_x: float, width_y: float, nterms: int = 1001) -> float:
    """Steady volumetric flow rate of a rectangular duct driven by a body force (no-slip walls):
    Q = F a b^3 ... written with half widths a (x) and b (y): Q = (4 F a^3 b / (3 nu)) (1 - 192 a / (pi^5 b)
    sum_{n odd} tanh(n pi b / (2a)) / n^5)  (White, Viscous Fluid Flow, eq. 3-48)."""
    a, b = 0.5 * width_x, 0.5 * width_y
    if a > b:
        a, b = b, a
    s = 0.0
    for n in range(1, nterms + 1, 2):
        s += math.tanh(n * math.pi * b / (2.0 * a)) / n ** 5
    return 4.0 * force * a ** 3 * b / (3.0 * nu) * (1.0 - 192.0 * a / (math.pi ** 5 * b) * s)


class SPSimulation:
    """One single-phase run on the GPU through libmflbm.so (mflbm_sp_* entry points).  Subclasses may bind
    another library with the same ABI."""

    prefix = "mflbm_"

    def _load(self):
        return native.load()

    def __init__(self, case: SPCase, state: Optional[SPState] = None):
        self.case = case
        self.lib = self._load()
        self.fn = native.bind(self.lib, self.prefix, single_phase=True)
        rp = case.rp
        self.nx, self.ny, self.nz = rp.nx, rp.ny, rp.nz
        self.ctx = C.c_void_p()
        rc = self.fn["create"](C.byref(case.params), C.byref(self.ctx))
        if rc != 0:
            msg = self.lib.mflbm_create_error().decode() if self.prefix == "mflbm_" else ""
            raise RuntimeError(f"{self.prefix}sp_create failed ({rc}): {msg}")
        self._call("upload_geometry", i8ptr(case.walls_h2), dptr(case.w_in))
        self.upload(state if state is not None else case.initial_state())
        self.monitor_prev = 0.0

    def _call(self, name: str, *args) -> None:
        rc = self.fn[name](self.ctx, *args)
        if rc != 0:
            msg = ""
            if "last_error" in self.fn:
                m = self.fn["last_error"](self.ctx)
                msg = m.decode() if m else ""
            raise RuntimeError(f"{self.prefix}sp_{name} failed ({rc}): {msg}")

    def close(self) -> None:
        if self.ctx:
            self.fn["destroy"](self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, st: SPState) -> None:
        self._keep = st
        self._call("upload_state", native.ptr_array(list(st.f)), dptr(np.ascontiguousarray(st.f_convec)))

    def set_equilibrium(self, rho: Optional[np.ndarray] = None, velocity: Optional[tuple] = None) -> None:
        """The state set on the device to SPCase.collision_equilibrium_state(rho, velocity) (halo-1 arrays;
        mflbm_sp_set_equilibrium): only the given fields cross PCIe."""
        args = []
        keep = []
        for a in [rho] + list(velocity if velocity is not None else (None, None, None)):
            if a is None:
                args.append(None)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                keep.append(a)
                args.append(dptr(a))
        self._call("set_equilibrium", args[0], args[1], args[2], args[3])

    def state(self) -> SPState:
        st = SPState.empty(self.nx, self.ny, self.nz)
        self._call("download_state", native.ptr_array(list(st.f)), dptr(st.f_convec))
        return st

    def step(self, ntime: int) -> None:
        self._call("step", ntime)

    def run(self, ntime0: int, nsteps: int) -> None:
        self._call("run", ntime0, nsteps)

    def synchronize(self) -> None:
        self._call("synchronize")

    def set_scalars(self, force_z: float, rho_in: float, uin_avg: float, relaxation: float = 1.0) -> None:
        self._call("set_scalars", force_z, rho_in, uin_avg, relaxation)

    def macro(self) -> Dict[str, np.ndarray]:
        self._call("compute_macro")
        sh = (self.nz + 2, self.ny + 2, self.nx + 2)
        u, v, w, rho = (np.zeros(sh) for _ in range(4))
        self._call("download_macro", dptr(u), dptr(v), dptr(w), dptr(rho))
        return {"u": u, "v": v, "w": w, "rho": rho}

    def monitor(self) -> SPMonitorResult:
        layers = np.zeros((2, self.nz))
        um = C.c_double()
        self._call("monitor", dptr(layers), C.byref(um))
        return sp_post_monitor(self.case, layers, um.value)

    def run_steady(self, ntime0: int, ntime_max: int, ntime_monitor: int,
                   tol: Optional[float] = None) -> tuple:
        """The reference's main loop without output files (Main.F90:124-230): step, and every ntime_monitor
        steps monitor; with steady_state_option 1 stop when |Q - Q_prev| / (|Q| + eps) < tol and
        ntime > ntime_monitor (Monitor.F90:145-156); stop on NaN or umax > 0.5 (:158-165).
        Returns (last step, status, list of (ntime, flowrate, rel_err)) with status 'converged', 'failed' or
        'max_steps'."""
        rp = self.case.rp
        tol = rp.convergence_criteria if tol is None else tol
        hist: List[tuple] = []
        ntime = ntime0
        while ntime <= ntime_max:
            # steps up to the next monitor step go to the device as one batch
            nm = -(-ntime // ntime_monitor) * ntime_monitor
            end = min(nm, ntime_max + 1)
            if end > ntime:
                self.run(ntime, end - ntime)
                ntime = end
            if ntime > ntime_max:
                break
            self.step(ntime)
            m = self.monitor()
            prev, self.monitor_prev = self.monitor_prev, m.flowrate
            rel = abs(m.flowrate - prev) / (abs(m.flowrate) + EPS)
            hist.append((ntime, m.flowrate, rel))
            if m.failed:
                return ntime, "failed", hist
            if rp.steady_state_option == 1 and rel < tol and ntime > ntime_monitor:
                return ntime, "converged", hist
            ntime += 1
        return ntime - 1, "max_steps", hist
