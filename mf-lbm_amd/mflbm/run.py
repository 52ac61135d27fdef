"""Drop-in run driver for a reference run directory: the main program of the two-phase solver
(multiphase_3D/0.src/Main_multiphase.F90:20-330) over this package's GPU path.

A run directory holds what the reference's ``irun.sh`` prepares (run_template/template-irun.sh:6-54):
``simulation_control.txt``, ``job_status.txt`` and ``path_info.txt`` (line 4 the wall file, line 6 the
boundary-info file).  ``run_directory`` reads them, builds the run (new, or continued from
``out2.checkpoint/id0000``), steps it, and writes what the reference writes:

* ``out1.output/``: ``info.txt`` (Init_multiphase.F90:78-95), the monitor files (``saturation.dat``,
  ``saturation_full_domain.dat``, ``Ca_number.dat``, ``flowrate_time.dat``, ``pre.dat``,
  ``profile/monitor*.dat``: Monitor.F90:153-248; ``steady_monitor_*.dat``: :265-268, :348, :432) and
  ``time.dat`` (Main_multiphase.F90:171-182), every line in the reference's edit descriptors;
* ``out2.checkpoint/id0000`` (and ``2rd_backup/``) on the wall-clock cadence and at the end
  (Main_multiphase.F90:223-259, :266-320; IO_multiphase.F90:562-642);
* ``out3.field_data/``: the phase and full-flow VTK files (Main_multiphase.F90:124-133, :198-215);
* ``job_status.txt``: the state machine's next state (Main_multiphase.F90:266-320).

The stop rules are the reference's ``simulation_end_indicator``: 0 maximum step, 1 converged (saturation,
capillary pressure or phase-field steady state, or breakthrough), 2 wall-clock limit, 3 failed (NaN or
largest velocity above 0.5) (Monitor.F90:253-272, :340-358, :435-456, :500-503).
"""
from __future__ import annotations

import dataclasses
import math
import os
import time
from typing import Callable, Dict, Optional, Type

import numpy as np

from . import io, vtk
from .case import Case, State
from .solver import MonitorResult, Simulation

EPS = 1.110223025e-16          # Module.F90:8

END_NAMES = {0: "simulation_reached_max_step", 1: "simulation_done", 2: "continue_simulation",
             3: "simulation_failed"}


class RunError(RuntimeError):
    """A condition on which the reference calls mpi_abort (missing file, finished job, bad status)."""


# -------------------------------------------------------------------------------------------------
# timers (IO_multiphase.F90:228-281, Init_multiphase.F90:154-183, :226-233)
# -------------------------------------------------------------------------------------------------

@dataclasses.dataclass
class Timers:
    ntime_max: int
    ntime_monitor: int
    ntime_monitor_profile: int
    ntime_animation: int
    ntime_visual: int
    ntime_clock_sum: int
    ntime_display: int


def _even_steps(x: float) -> int:
    """An injected-volume cadence in steps: truncated to an integer, odd values moved to the next even step
    (the populations are in natural order only after even steps)."""
    n = int(x)
    return n + 1 if n % 2 == 1 else n


def derive_timers(ctl: Dict[str, object], rp: io.RunParams, pore_sum: int) -> Timers:
    c = dict(io.DEFAULT_CONTROL)
    c.update(ctl)
    t = Timers(int(c["max_time_step"]), int(c["monitor_timer"]), 0, int(c["ntime_animation"]),
               int(c["ntime_visual"]), int(c["computation_time_timer"]), int(c["display_steps_timer"]))
    t.ntime_monitor_profile = int(c["monitor_profile_timer_ratio"]) * t.ntime_monitor
    if rp.open_z and rp.inlet_bc == 1:
        uin_avg_0 = float(c["capillary_number"]) * rp.gamma / rp.la_nu1
        flowrate = uin_avg_0 * rp.A_xy
        if float(c["target_inject_pore_volume"]) > 0:
            t.ntime_max = _even_steps(float(c["target_inject_pore_volume"]) * pore_sum / flowrate)
        if float(c["d_vol_animation"]) > 0:
            t.ntime_animation = _even_steps(float(c["d_vol_animation"]) * pore_sum / flowrate)
        if float(c["d_vol_detail"]) > 0:
            t.ntime_visual = _even_steps(float(c["d_vol_detail"]) * pore_sum / flowrate)
        if float(c["d_vol_monitor"]) > 0:
            t.ntime_monitor = _even_steps(float(c["d_vol_monitor"]) * pore_sum / flowrate)
            t.ntime_monitor_profile = int(c["monitor_profile_timer_ratio"]) * t.ntime_monitor
    return t


# -------------------------------------------------------------------------------------------------
# monitor output and stop rules
# -------------------------------------------------------------------------------------------------

class Monitors:
    """The monitor variants of the main loop (Main_multiphase.F90:184-196) with their files and the state the
    reference keeps between calls (monitor_previous_value / monitor_current_value, Init_multiphase.F90:186-187)."""

    def __init__(self, out1: str, case: Case, ctl: Dict[str, object], timers: Timers):
        c = dict(io.DEFAULT_CONTROL)
        c.update(ctl)
        self.out1 = out1
        self.case = case
        self.timers = timers
        self.option = int(c["steady_state_option"])
        self.breakthrough = int(c["breakthrough_check"]) == 1
        self.criterion = float(c["convergence_criteria"])
        self.ntime_relaxation = 0   # never set from the control file: the module variable keeps 0
        self.previous = 0.0
        self.current = 0.0
        # The capillary-pressure monitor never resets its local umax before the max reduction (Monitor.F90:378-392);
        # with the reference build it keeps the previous call's value, so it reports the running maximum over
        # the run's monitor calls (starting from 0)
        self.pc_umax = 0.0
        self.pore = (case.walls_global <= 0).sum(axis=(1, 2)).astype(np.int64)
        self.messages = []

    def _append(self, name: str, line: str) -> None:
        with open(os.path.join(self.out1, name), "a") as fh:
            fh.write(line + "\n")

    def __call__(self, sim: Simulation, ntime: int) -> int:
        """Run the monitor of this step; returns the end indicator it sets (0 = keep going)."""
        if self.option in (0, 3):
            m = sim.monitor()
            end = self._unsteady(m, ntime)
            if self.breakthrough:
                end = self._breakthrough(m, end)
            return end
        if self.option == 1:
            return self._capillary_pressure(sim.monitor(), ntime)
        if self.option == 2:
            return self._phase_field(sim.monitor(update_phi_old=True), ntime)
        return 0

    # monitor, Monitor.F90:5-277
    def _unsteady(self, m: MonitorResult, ntime: int) -> int:
        rp = self.case.rp
        nz = rp.nz
        fl1, fl2, vol1, vol2, mass1, mass2, pre = m.layers
        k0, k1 = rp.n_exclude_inlet, nz - rp.n_exclude_outlet
        self._append("saturation.dat", io.fortran_line(ntime, "e14.7", m.saturation, m.vol1_sum, m.vol2_sum,
                                                       m.mass1_sum, m.mass2_sum))
        t1 = t2 = t3 = t4 = 0.0
        for k in range(nz):
            t1 += mass1[k]
            t2 += mass2[k]
            t3 += vol1[k]
            t4 += vol2[k]
        f1w = f2w = 0.0
        for k in range(nz):
            f1w += fl1[k]
            f2w += fl2[k]
        f1w, f2w = f1w / nz, f2w / nz
        nlay = nz - rp.n_exclude_outlet - rp.n_exclude_inlet
        f1 = f2 = 0.0
        for k in range(k0, k1):
            f1 += fl1[k]
            f2 += fl2[k]
        f1, f2 = f1 / nlay, f2 / nlay
        self._append("Ca_number.dat", io.fortran_line(ntime, "e14.7", m.ca, m.umax, *m.kinetic_energy))
        self._append("flowrate_time.dat", io.fortran_line(ntime, "E14.6", f1w + f2w, f1w, f2w, f1 + f2, f1, f2))
        self._append("saturation_full_domain.dat", io.fortran_line(ntime, "e14.7", m.saturation_full_domain,
                                                                   t3, t4, t1, t2))
        if rp.open_z:
            p1 = pre[k0] / self.pore[k0]
            p2 = pre[k1 - 1] / self.pore[k1 - 1]
            self._append("pre.dat", io.fortran_line(ntime, "e14.7", p1, p2, p1 - p2))
        if ntime % self.timers.ntime_monitor_profile == 0:
            os.makedirs(os.path.join(self.out1, "profile"), exist_ok=True)
            rows = []
            for k in range(nz):
                rows.append(io.fortran_fields("e14.7", float(k + 1), fl1[k], fl2[k], mass1[k] / (mass1[k] + mass2[k]),
                                              pre[k] / (float(self.pore[k]) + EPS)))
            with open(os.path.join(self.out1, "profile", "monitor" + io.profile_name(ntime)), "w") as fh:
                fh.write("".join(r + "\n" for r in rows))
        end = 0
        if math.isnan(m.saturation_full_domain) or math.isnan(m.ca):
            self.messages.append('Simulation failed due to NAN of "saturation or capillary number"!')
            end = 3
        elif m.umax > 0.5:
            self.messages.append("Simulation failed due to maximum velocity larger than 0.5!")
            end = 3
        if self.option == 3:
            self.previous, self.current = self.current, m.saturation_full_domain
            err = abs(self.current - self.previous) / (abs(self.current) + EPS)
            self._append("steady_monitor_saturation_error.dat", io.fortran_line(ntime, "e14.7", err))
            if err < self.criterion and ntime > self.timers.ntime_monitor:
                self.messages.append(f"Simulation converged based on saturation! Relative error = {err}")
                end = 1
        return end

    # monitor_breakthrough, Monitor.F90:472-507 (fluid 1 at any fluid node of layer nz - 1)
    def _breakthrough(self, m: MonitorResult, end: int) -> int:
        if m.outlet_phase1 >= 1:
            self.messages.append("Breakthrough point reached! Exiting program!")
            return 1
        return end

    # monitor_multiphase_steady_capillarypressure, Monitor.F90:370-462
    def _capillary_pressure(self, m: MonitorResult, ntime: int) -> int:
        pc = m.capillary_pressure
        if not (m.umax <= self.pc_umax):
            self.pc_umax = m.umax
        umax = self.pc_umax
        self.previous, self.current = self.current, pc
        err = abs(self.current - self.previous) / (abs(self.current) + EPS)
        self._append("steady_monitor_capillary_pressure_error.dat",
                     io.fortran_line(ntime, "e14.7", err, pc, m.pre_w, m.pre_nw, err, umax))
        if math.isnan(umax) or math.isnan(pc):
            self.messages.append("Simulation failed due to NAN!")
            return 3
        if umax < 0.5:
            if err < self.criterion and ntime > self.timers.ntime_monitor:
                self.messages.append(f"Simulation converged based on change of capillary pressure! Relative error = {err}")
                return 1
            return 0
        self.messages.append("maximum velocity larger than 0.5, simulation failed!!!")
        return 3

    # monitor_multiphase_steady_phasefield, Monitor.F90:287-365
    def _phase_field(self, m: MonitorResult, ntime: int) -> int:
        if ntime <= self.ntime_relaxation:
            return 0
        if math.isnan(m.umax) or math.isnan(m.dphi_max):
            self.messages.append("Simulation failed due to NAN!")
            return 3
        self._append("steady_monitor_max_phi_change.dat", io.fortran_line(ntime, "e14.7", m.dphi_max, m.umax))
        if m.umax < 0.5:
            if m.dphi_max < self.criterion and ntime > self.timers.ntime_monitor:
                self.messages.append(f"Simulation converged based on maximum local change of phi! Relative error = {m.dphi_max}")
                return 1
            return 0
        self.messages.append("Maximum velocity larger than 0.5, simulation failed!!!")
        return 3


# -------------------------------------------------------------------------------------------------
# the run
# -------------------------------------------------------------------------------------------------

@dataclasses.dataclass
class RunResult:
    end_indicator: int
    ntime: int                 # last step taken (the checkpoint holds ntime + 1 as its next step)
    job_status: str
    steps: int                 # steps taken in this invocation
    seconds: float
    messages: list


def read_path_info(run_dir: str):
    """(wall file, boundary-info file) from path_info.txt: lines 4 and 6 (Misc.F90:50-59)."""
    p = os.path.join(run_dir, "path_info.txt")
    if not os.path.exists(p):
        raise RunError("Error! path_info.txt is not found! Exiting program!")
    with open(p) as fh:
        lines = fh.read().splitlines()
    if len(lines) < 4:
        raise RunError("path_info.txt: the geometry file path is missing")
    geo = lines[3].strip()
    bnd = lines[5].strip() if len(lines) >= 6 else ""
    return geo, bnd


def read_job_status(run_dir: str) -> str:
    """The job status and the reference's reaction to it (IO_multiphase.F90:17-54)."""
    p = os.path.join(run_dir, "job_status.txt")
    if not os.path.exists(p):
        raise RunError("Missing job status file! Exiting program!")
    with open(p) as fh:
        status = fh.read().split()
    status = status[0] if status else ""
    if status in ("new_simulation", "continue_simulation"):
        return status
    msg = {"simulation_done": "Previous simulation is already finished! Exiting program!",
           "simulation_reached_max_step": "Previous simulation reached maximum step! Exiting program!",
           "simulation_failed": "Previous simulation failed! Exiting program!"}.get(status, "Wrong simlation status! Exiting program!")
    raise RunError(msg)


def boundary_counts(walls_global: np.ndarray, jper: int, kper: int) -> tuple:
    """num_solid_boundary_global and num_fluid_boundary_global of geometry_preprocessing_new
    (Geometry_preprocessing.F90:47-184): the wall field is extended by 10 ghost layers (replicated, or wrapped
    along periodic y/z; z first, then y, then x), solid nodes with a non-solid neighbour and non-solid nodes with
    a solid neighbour are counted over the domain and 4 ghost layers."""
    gl = 10
    t = np.pad(walls_global.astype(np.int8), ((gl, gl), (0, 0), (0, 0)), mode="wrap" if kper else "edge")
    t = np.pad(t, ((0, 0), (gl, gl), (0, 0)), mode="wrap" if jper else "edge")
    t = np.pad(t, ((0, 0), (0, 0), (gl, gl)), mode="edge")
    core = (slice(1, -1),) * 3
    any_open = np.zeros(t[core].shape, bool)
    any_solid = np.zeros(t[core].shape, bool)
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
                       (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1),
                       (0, 1, -1), (0, -1, 1)):
        nb = t[1 + dz:t.shape[0] - 1 + dz, 1 + dy:t.shape[1] - 1 + dy, 1 + dx:t.shape[2] - 1 + dx]
        any_open |= nb <= 0
        any_solid |= nb >= 1
    c = t[core]
    win = (slice(gl - 5, -(gl - 5)),) * 3     # nodes 1 - 4 .. n + 4 of the core (which starts at node 2 - gl)
    return int(((c == 1) & any_open)[win].sum()), int(((c == 0) & any_solid)[win].sum())


def write_info(path: str, case: Case, ns_global: int, nf_global: int) -> None:
    """out1.output/info.txt (Init_multiphase.F90:78-95)."""
    rp = case.rp
    nx, ny, nz = rp.nx, rp.ny, rp.nz
    pore = (case.walls_global <= 0).sum(axis=(1, 2))
    pore_sum = int(pore.sum())
    pore_eff = int(pore[rp.n_exclude_inlet:nz - rp.n_exclude_outlet].sum())
    full = pore_sum / float((nx - 2) * (ny - 2) * nz)
    eff = pore_eff / float((nx - 2) * (ny - 2)) / float(nz - rp.n_exclude_outlet - rp.n_exclude_inlet)
    lines = [" Grid information:", f"nxGlobal = {nx:4d}, nyGlobal = {ny:4d}, nzGlobal = {nz:4d}",
             f"Inlet open cross sectional area = {io.fortran_f(rp.A_xy, 14, 2)}", "  Pore information:",
             f"Total number of pore nodes = {pore_sum:14d}", f"Total number of solid boundary nodes = {ns_global:14d}",
             f"Total number of fluid boundary nodes = {nf_global:14d}",
             f"Total number of effective pore nodes (excluding inlet/outlet) = {pore_eff:14d}",
             f"Full domain porosity = {io.fortran_f(full, 6, 4)}",
             f"Effective domain porosity  (excluding inlet/outlet) = {io.fortran_f(eff, 6, 4)}"]
    with open(path, "w") as fh:
        fh.write("".join(s + "\n" for s in lines))


def _checkpoint(sim: Simulation, ntime: int, path: str, convective: bool) -> None:
    """save_checkpoint (IO_multiphase.F90:562-642): next step, force_z, rho_in, then the state."""
    st = sim.state()
    rp = sim.case.rp
    p = sim.case.params
    ck = io.Checkpoint(ntime + 1, float(p.force_z), float(p.rho_in), st.f, st.g, st.phi)
    if convective:
        ck.f_convec, ck.g_convec, ck.phi_convec = st.f_convec, st.g_convec, st.phi_convec
    io.write_checkpoint(path, ck)
    _ = rp


def _vtk(sim: Simulation, case: Case, ntime: int, kind: int, out3: str, double: bool) -> None:
    """VTK_legacy_writer_3D type 1 (full flow field) or 2 (phase distribution), IO_multiphase.F90:824-1003."""
    m = sim.macro()
    phi = m["phi"][4:-4, 4:-4, 4:-4]
    if kind == 2:
        vtk.write_phase(os.path.join(out3, "phase_distribution", "small_" + vtk.vtk_name(ntime)), phi, case.walls_global)
    else:
        s = (slice(1, -1),) * 3
        vtk.write_full_flow(os.path.join(out3, "full_flow_field", "full_flow_field_" + vtk.vtk_name(ntime)), phi,
                            m["rho"][s], m["u"][s], m["v"][s], m["w"][s], double=double)


def run_directory(run_dir: str, backend: Type[Simulation] = Simulation, precision: str = "fp64",
                  max_steps: Optional[int] = None, clock: Callable[[], float] = time.monotonic,
                  log: Callable[[str], None] = lambda s: None) -> RunResult:
    """Run (or continue) the run in ``run_dir`` to its end condition.  ``backend`` is the class that drives
    the lattice update (the GPU path by default).  The wall-clock timers of the control file
    (checkpoint_save_timer, checkpoint_2rd_save_timer, simulation_duration_timer, in hours) are checked every
    computation_time_timer steps as the reference does; ``max_steps`` additionally ends this invocation after
    that many steps the way the duration timer does (checkpoint, status continue_simulation)."""
    t_start = clock()
    status = read_job_status(run_dir)
    ctl_path = os.path.join(run_dir, "simulation_control.txt")
    if not os.path.exists(ctl_path):
        raise RunError("Missing simulation_control.txt! Exiting program!")
    ctl = io.read_control(ctl_path)
    c = dict(io.DEFAULT_CONTROL)
    c.update(ctl)
    if int(c["benchmark_cmd"]) == 1:
        raise RunError("benchmark_cmd 1: use bench.py (the benchmark loop) instead of the run driver")
    if int(c["extreme_large_sim_cmd"]) == 1:
        raise RunError("extreme_large_sim_cmd 1 (per-rank raw dumps) is not supported by the run driver")
    if int(c["external_geometry_read_cmd"]) != 1:
        raise RunError("external_geometry_read_cmd must be 1: the geometry comes from the wall file")
    geo, bnd = read_path_info(run_dir)
    geo = geo if os.path.isabs(geo) else os.path.join(run_dir, geo)
    if not os.path.exists(geo):
        raise RunError("Error! No external geometry file found! Exiting program!")
    sample = io.read_walls(geo)
    boundary = None
    if int(c["geometry_preprocess_cmd"]) == 1:
        boundary = bnd if os.path.isabs(bnd) else os.path.join(run_dir, bnd)
        if not os.path.exists(boundary):
            raise RunError("Error! No boundary info file found! Exiting program!")
    case = Case(ctl, sample, boundary_file=boundary, precision=precision)
    rp = case.rp
    pore_sum = case.n_fluid
    timers = derive_timers(ctl, rp, pore_sum)
    out1 = os.path.join(run_dir, "out1.output")
    out2 = os.path.join(run_dir, "out2.checkpoint")
    out3 = os.path.join(run_dir, "out3.field_data")
    for d in (out1, os.path.join(out1, "profile"), out2, os.path.join(out2, "2rd_backup"),
              os.path.join(out3, "phase_distribution"), os.path.join(out3, "full_flow_field")):
        os.makedirs(d, exist_ok=True)
    write_info(os.path.join(out1, "info.txt"), case, *boundary_counts(case.walls_global, rp.jper, rp.kper))
    convective = rp.outlet_bc == 1
    if status == "new_simulation":
        ntime0 = 1
        sim = backend(case, state=case.initial_state())
    else:
        ck_path = os.path.join(out2, "id0000")
        if not os.path.exists(ck_path):
            raise RunError("Checkpoint data not found! Exiting program!")
        ck = io.read_checkpoint(ck_path, rp.nx, rp.ny, rp.nz, convective)
        ntime0 = ck.ntime_next
        case.params.force_z = ck.force_z
        case.params.rho_in = ck.rho_in
        sim = backend(case, state=State.from_checkpoint(ck))
    double = int(c["output_fieldData_precision_cmd"]) == 1
    mon = Monitors(out1, case, ctl, timers)
    hours = lambda: (clock() - t_start) / 3600.0
    ck_timer, ck2_timer = float(c["checkpoint_save_timer"]), float(c["checkpoint_2rd_save_timer"])
    duration = float(c["simulation_duration_timer"])
    n_ck = n_ck2 = 1
    try:
        if status == "new_simulation":
            if mon.option == 2:
                sim.monitor(update_phi_old=True)   # phi_old = phi of the initial state (Init_multiphase.F90:341-351)
            _vtk(sim, case, 0, 2, out3, double)
            vtk.write_walls(os.path.join(out3, "walls_bin.vtk"), case.walls_global)
        end = 0
        ntime = ntime0
        t_sum = 0.0
        taken = 0
        for ntime in range(ntime0, timers.ntime_max + 1):
            t0 = clock()
            sim.step(ntime)
            taken += 1
            if ntime % timers.ntime_clock_sum == 0:
                # time of the steps since the last entry (the reference sums per-step wall times)
                sim.synchronize()
                with open(os.path.join(out1, "time.dat"), "a") as fh:
                    fh.write(f"{ntime:10d} {io.fortran_f(t_sum, 10, 2)} sec\n")
                t_sum = 0.0
            if ntime % timers.ntime_monitor == 0:
                end = mon(sim, ntime)
            if ntime % timers.ntime_display == 0:
                log(f"ntime = {ntime}")
            if ntime % timers.ntime_animation == 0:
                _vtk(sim, case, ntime, 2, out3, double)
            if ntime % timers.ntime_visual == 0:
                _vtk(sim, case, ntime, 1, out3, double)
            if ntime % timers.ntime_clock_sum == 0:   # Main_multiphase.F90:223-259
                h = hours()
                if h >= duration:
                    mon.messages.append("Time to save checkpoint data and exit the program!")
                    end = 2
                if h >= n_ck * ck_timer:
                    _checkpoint(sim, ntime, os.path.join(out2, "id0000"), convective)
                    with open(os.path.join(run_dir, "job_status.txt"), "w") as fh:
                        fh.write("continue_simulation\n")
                    n_ck += 1
                if h >= n_ck2 * ck2_timer:
                    _checkpoint(sim, ntime, os.path.join(out2, "2rd_backup", "id0000"), convective)
                    n_ck2 += 1
            if end == 0 and max_steps is not None and taken >= max_steps:
                end = 2
            t_sum += clock() - t0
            if end > 0:
                break
        else:
            ntime = timers.ntime_max   # the loop ran out: the reference dials ntime back to ntime_max
        if end in (0, 1, 2):
            _checkpoint(sim, ntime, os.path.join(out2, "id0000"), convective)
        if end in (0, 1):
            _vtk(sim, case, ntime, 2, out3, double)
        if end == 3:
            _vtk(sim, case, ntime, 1, out3, double)
        job = END_NAMES[end]
        with open(os.path.join(run_dir, "job_status.txt"), "w") as fh:
            fh.write(job + "\n")
        for msg in mon.messages:
            log(msg)
        return RunResult(end, ntime, job, taken, clock() - t_start, list(mon.messages))
    finally:
        sim.close()
