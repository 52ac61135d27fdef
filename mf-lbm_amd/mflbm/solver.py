"""Host-side driver of one run on the GPU, mirroring the operations the reference's main loop performs on
its module state (Main_multiphase.F90:104-264): upload once, step, monitor, download."""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Dict, Optional

import numpy as np

from . import native
from .case import Case, State
from .native import dptr, i8ptr, iptr

EPS = 1.110223025e-16


@dataclasses.dataclass
class MonitorResult:
    """Rank-0 quantities of monitor / cal_saturation (Monitor.F90:153-272, :512-550)."""
    saturation: float
    saturation_full_domain: float
    ca: float
    umax: float
    kinetic_energy: tuple
    fl_avg_whole: float
    fl_avg: float
    vol1_sum: float
    vol2_sum: float
    mass1_sum: float
    mass2_sum: float
    pressure_drop: Optional[float]
    capillary_pressure: float
    pre_w: float                # mean density of the wetting (phi < -0.99) and non-wetting (phi > 0.99) nodes
    pre_nw: float
    dphi_max: float
    outlet_phase1: int
    failed: bool
    layers: np.ndarray          # (7, nz): fl1, fl2, vol1, vol2, mass1, mass2, pre


def post_monitor(case: Case, layers: np.ndarray, s) -> MonitorResult:
    rp = case.rp
    nz = rp.nz
    fl1, fl2, vol1, vol2, mass1, mass2, pre = layers
    k0, k1 = rp.n_exclude_inlet, nz - rp.n_exclude_outlet
    mass1_sum = mass2_sum = vol1_sum = vol2_sum = 0.0
    for k in range(k0, k1):
        mass1_sum += mass1[k]
        mass2_sum += mass2[k]
        vol1_sum += vol1[k]
        vol2_sum += vol2[k]
    saturation = vol1_sum / (vol1_sum + vol2_sum) if (vol1_sum + vol2_sum) != 0 else float("nan")
    t3 = t4 = 0.0
    for k in range(nz):
        t3 += vol1[k]
        t4 += vol2[k]
    sat_full = t3 / (t3 + t4) if (t3 + t4) != 0 else float("nan")
    f1w = f2w = 0.0
    for k in range(nz):
        f1w += fl1[k]
        f2w += fl2[k]
    fl_whole = f1w / nz + f2w / nz
    f1 = f2 = 0.0
    for k in range(k0, k1):
        f1 += fl1[k]
        f2 += fl2[k]
    nlay = nz - rp.n_exclude_outlet - rp.n_exclude_inlet
    fl_avg = f1 / nlay + f2 / nlay
    ca = (fl_avg / rp.A_xy) * rp.la_nu1 / rp.gamma
    umax = math.sqrt(s.umax2)
    dp = None
    if rp.open_z:
        pore = (case.walls_global <= 0).sum(axis=(1, 2))
        t1 = pre[k0] / pore[k0] if pore[k0] else float("nan")
        t2 = pre[k1 - 1] / pore[k1 - 1] if pore[k1 - 1] else float("nan")
        dp = t1 - t2
    pre_w = s.pre_w / (s.n_w + EPS)
    pre_nw = s.pre_nw / (s.n_nw + EPS)
    failed = math.isnan(sat_full) or math.isnan(ca) or umax > 0.5
    return MonitorResult(saturation, sat_full, ca, umax, (0.5 * s.usq1, 0.5 * s.usq2), fl_whole, fl_avg, vol1_sum,
                         vol2_sum, mass1_sum, mass2_sum, dp, (pre_nw - pre_w) / 3.0, pre_w, pre_nw, s.dphi_max,
                         int(s.outlet_phase1), failed, layers)


class Simulation:
    """One run on the GPU through libmflbm.so.  Subclasses may bind another library with the same ABI."""

    prefix = "mflbm_"

    def _load(self):
        return native.load()

    def __init__(self, case: Case, state: Optional[State] = None, gradient: bool = True,
                 device_init: bool = False):
        """``state`` is uploaded when given; otherwise the new-run initial state is built on the host and
        uploaded, or with ``device_init`` computed on the device (mflbm_init_state_device, at rest)."""
        self.case = case
        self.lib = self._load()
        self.fn = native.bind(self.lib, self.prefix)
        rp = case.rp
        self.nx, self.ny, self.nz = rp.nx, rp.ny, rp.nz
        self.ctx = C.c_void_p()
        rc = self.fn["create"](C.byref(case.params), C.byref(self.ctx))
        if rc != 0:
            msg = self.lib.mflbm_create_error().decode() if self.prefix == "mflbm_" else ""
            raise RuntimeError(f"{self.prefix}create failed ({rc}): {msg}")
        b = case.lists
        self._call("upload_geometry", i8ptr(case.walls_h2), len(b.s_nfl), iptr(b.s_ijk), iptr(b.s_nfl),
                   iptr(b.s_nbr), dptr(b.s_w), len(b.f_theta), iptr(b.f_ijk), dptr(b.f_nw), dptr(b.f_theta),
                   dptr(case.w_in))
        if state is not None or not device_init:
            self.upload(state if state is not None else case.initial_state())
        else:
            self._call("init_state_device", case.rp.init_option, case.rp.interface_z0)
        if gradient:
            self.color_gradient()

    # -- plumbing --------------------------------------------------------------------------------------
    def _call(self, name: str, *args) -> None:
        rc = self.fn[name](self.ctx, *args)
        if rc != 0:
            msg = ""
            if "last_error" in self.fn:
                m = self.fn["last_error"](self.ctx)
                msg = m.decode() if m else ""
            raise RuntimeError(f"{self.prefix}{name} failed ({rc}): {msg}")

    def close(self) -> None:
        if self.ctx:
            self.fn["destroy"](self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -----------------------------------------------------------------------------------------
    def upload(self, st: State) -> None:
        self._keep = st
        self._call("upload_state", native.ptr_array(list(st.f)), native.ptr_array(list(st.g)), dptr(st.phi),
                   dptr(np.ascontiguousarray(st.f_convec)), dptr(np.ascontiguousarray(st.g_convec)),
                   dptr(np.ascontiguousarray(st.phi_convec)))

    def state(self) -> State:
        st = State.empty(self.nx, self.ny, self.nz)
        self._call("download_state", native.ptr_array(list(st.f)), native.ptr_array(list(st.g)), dptr(st.phi),
                   dptr(st.f_convec), dptr(st.g_convec), dptr(st.phi_convec))
        return st

    # -- hot path --------------------------------------------------------------------------------------
    def color_gradient(self) -> None:
        self._call("color_gradient")

    def step(self, ntime: int) -> None:
        self._call("step", ntime)

    def run(self, ntime0: int, nsteps: int) -> None:
        self._call("run", ntime0, nsteps)

    def synchronize(self) -> None:
        if "synchronize" in self.fn:   # a host-side binding of the same ABI has nothing to wait for
            self._call("synchronize")

    # -- diagnostics -----------------------------------------------------------------------------------
    def macro(self) -> Dict[str, np.ndarray]:
        self._call("compute_macro")
        sh = (self.nz + 2, self.ny + 2, self.nx + 2)
        u, v, w, rho = (np.zeros(sh) for _ in range(4))
        phi = np.zeros((self.nz + 8, self.ny + 8, self.nx + 8))
        self._call("download_macro", dptr(u), dptr(v), dptr(w), dptr(rho), dptr(phi))
        return {"u": u, "v": v, "w": w, "rho": rho, "phi": phi}

    def gradient(self) -> Dict[str, np.ndarray]:
        s2 = (self.nz + 4, self.ny + 4, self.nx + 4)
        cx, cy, cz, cn = (np.zeros(s2) for _ in range(4))
        curv = np.zeros((self.nz + 2, self.ny + 2, self.nx + 2))
        self._call("download_gradient", dptr(cx), dptr(cy), dptr(cz), dptr(cn), dptr(curv))
        return {"cn_x": cx, "cn_y": cy, "cn_z": cz, "c_norm": cn, "curv": curv}

    def monitor(self, update_phi_old: bool = False) -> MonitorResult:
        layers = np.zeros((7, self.nz))
        s = native.MonitorSums()
        self._call("monitor", dptr(layers), C.byref(s), 1 if update_phi_old else 0)
        return post_monitor(self.case, layers, s)
