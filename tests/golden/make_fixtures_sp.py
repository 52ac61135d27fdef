#!/usr/bin/env python3
"""Generate the single-phase golden fixtures from the reference's single-phase solver built by
oracle/build_ref.sh (oracle/_ref/MF_LBM_singlephase.cpu, singleton MPI process).

The unmodified binary runs small synthetic cases; its checkpoint after ``max_time_step`` = N steps holds the
exact state (19 populations with their halo and the convective-outlet history; save_checkpoint,
singlephase_3D/0.src/IO.F90:490-547, written after the loop dials ntime back, Main.F90:222-227).

Output: tests/golden/sp_<case>.npz with
  * ``sha_<n>_f``, ``sha_<n>_f_int`` (and ``sha_<n>_convec`` with the convective outlet): sha256 of the
    exact little-endian bytes of the state after n steps
  * ``f_<N>``: interior populations after the last step N
  * ``walls``, ``control`` (JSON of the control values)
  * ``monitor_flowrate_time`` / ``monitor_pre``: the monitor files of a run with the monitor on
    (Monitor.F90:119-130; E14.6 / E14.7 text).
The reference is used only through the built binary; this script is never run on a machine without it.
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(REPO, "mf-lbm_amd"))

from mflbm import geometry, io  # noqa: E402

REF_BIN = os.path.join(REPO, "oracle", "_ref", "MF_LBM_singlephase.cpu")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


CASES = {
    # z-periodic duct with the program's built-in off-axis sphere, body force, bounce-back-optimised MRT
    "sp_periodic_obstacle": dict(
        geom=lambda: geometry.duct_obstacle(14, 14, 24),
        control={"lattice_dimensions": (14, 14, 24), "periodic_indicator": (0, 0, 1),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 4),
                 "fluid_viscosity": 0.1, "body_force_0": 1e-5, "MRT_collision_parameter_preset": 1},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # velocity inlet (rectangular-duct profile) + convective outlet through a small sphere pack
    "sp_open_velocity": dict(
        geom=lambda: geometry.sphere_pack(12, 12, 28, 0.45, 2.0, 3.5, 7, buffer=4),
        control={"lattice_dimensions": (12, 12, 28), "periodic_indicator": (0, 0, 0),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 0),
                 "inlet_BC": 1, "outlet_BC": 1, "fluid_viscosity": 1.0 / 15.0, "Reynolds_number": 2.0,
                 "char_length": 10.0, "body_force_0": 0.0, "MRT_collision_parameter_preset": 1,
                 "excluded_layers": (4, 4)},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # Zou-He pressure inlet and outlet around a sphere in a tube, single-relaxation-time rates
    "sp_open_pressure": dict(
        geom=lambda: geometry.tube_sphere(12, 12, 24, 4.8, 2.6, buffer=3),
        control={"lattice_dimensions": (12, 12, 24), "periodic_indicator": (0, 0, 0),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 0),
                 "inlet_BC": 2, "outlet_BC": 2, "fluid_viscosity": 0.05, "rho_drop": 0.01,
                 "body_force_0": 0.0, "MRT_collision_parameter_preset": 0, "excluded_layers": (2, 2)},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # y- and z-periodic box around the obstacle (y faces and x-parallel edges of the exchange)
    "sp_periodic_yz": dict(
        geom=lambda: geometry.duct_obstacle(12, 14, 16),
        control={"lattice_dimensions": (12, 14, 16), "periodic_indicator": (0, 1, 1),
                 "domain_wall_status_y": (0, 0), "domain_wall_status_z": (0, 0),
                 "MPI_async_layers_num": (0, 4, 4), "fluid_viscosity": 1.0 / 6.0, "body_force_0": 2e-5,
                 "MRT_collision_parameter_preset": 1},
        steps=[0, 1, 2, 3, 40], monitor=10),
}


def run_ref(workdir: str, walls: np.ndarray, control: dict, nsteps: int, threads: int,
            monitor: int | None = None) -> float:
    os.makedirs(workdir, exist_ok=True)
    io.write_walls(os.path.join(workdir, "walls.dat"), walls)
    ctl = dict(control)
    ctl["max_time_step"] = nsteps
    if monitor:
        ctl["monitor_timer"] = monitor
        ctl["monitor_profile_timer_ratio"] = 1
    io.write_control_sp(os.path.join(workdir, "simulation_control.txt"), ctl)
    with open(os.path.join(workdir, "job_status.txt"), "w") as fh:
        fh.write("new_simulation\n")
    with open(os.path.join(workdir, "path_info.txt"), "w") as fh:
        fh.write("# executable location\n" + REF_BIN + "\n# geometry file\n" + os.path.join(workdir, "walls.dat") + "\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.time()
    r = subprocess.run([REF_BIN], cwd=workdir, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    dt = time.time() - t0
    if r.returncode != 0 or not os.path.exists(os.path.join(workdir, "out2.checkpoint", "id0000")):
        raise RuntimeError(f"reference run failed in {workdir}:\n{r.stdout[-3000:]}")
    return dt


def harvest(name: str, spec: dict, threads: int, keep: str | None) -> dict:
    walls = spec["geom"]()
    control = spec["control"]
    nx, ny, nz = control["lattice_dimensions"]
    conv = control.get("outlet_BC", 0) == 1
    out: dict = {"walls": walls, "control": json.dumps(control, sort_keys=True)}
    root = keep or tempfile.mkdtemp(prefix=f"mflbm_fix_{name}_")
    timings = {}
    for n in spec["steps"]:
        wd = os.path.join(root, f"step{n}")
        timings[n] = run_ref(wd, walls, control, n, threads)
        ck = io.read_checkpoint_sp(os.path.join(wd, "out2.checkpoint", "id0000"), nx, ny, nz, conv)
        if ck.ntime_next != n + 1:
            raise RuntimeError(f"{name}: checkpoint after {n} steps says next step {ck.ntime_next}")
        out[f"sha_{n}_f"] = sha(ck.f)
        out[f"sha_{n}_f_int"] = sha(ck.f[:, 1:-1, 1:-1, 1:-1])
        if conv:
            out[f"sha_{n}_convec"] = sha(ck.f_convec)
        if n == spec["steps"][-1]:
            out[f"f_{n}"] = ck.f[:, 1:-1, 1:-1, 1:-1].copy()
    if spec.get("monitor"):
        m = spec["monitor"]
        n = spec["steps"][-1]
        wd = os.path.join(root, "monitor")
        run_ref(wd, walls, control, n, threads, monitor=m)
        for fname in ("flowrate_time.dat", "pre.dat"):
            p = os.path.join(wd, "out1.output", fname)
            if os.path.exists(p):
                with open(p) as fh:
                    out["monitor_" + fname.replace(".dat", "")] = fh.read()
    out["ref_seconds"] = json.dumps(timings)
    out["ref_threads"] = threads
    if keep is None:
        shutil.rmtree(root, ignore_errors=True)
    return out


def main() -> None:
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--cases", nargs="*", default=None)
    ap.add_argument("--keep", default=None, help="keep run directories under this path")
    args = ap.parse_args()
    if not os.path.exists(REF_BIN):
        sys.exit("reference binary missing: run oracle/build_ref.sh first")
    todo = {k: v for k, v in CASES.items() if not args.cases or k in args.cases}
    for name, spec in todo.items():
        t = time.time()
        keep = os.path.join(args.keep, name) if args.keep else None
        data = harvest(name, spec, args.threads, keep)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
        print(f"{name}: {time.time() - t:.1f}s", flush=True)


if __name__ == "__main__":
    main()
