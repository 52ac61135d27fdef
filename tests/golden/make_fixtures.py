#!/usr/bin/env python3
"""Generate the golden fixtures from the reference solver built by oracle/build_ref.sh.

Runs the unmodified reference binary (oracle/_ref/MF_LBM_multiphase.cpu, singleton MPI process,
OMP_NUM_THREADS as given) on small synthetic cases and harvests its checkpoint files, which hold the
exact solver state (all 38 populations with halo, phi with its 4-layer halo and the convective-outlet
history).  A checkpoint written after ``max_time_step`` = N is the state after N time steps
(Main_multiphase.F90:264-270 dials ntime back and calls save_checkpoint).

Output: tests/golden/<case>.npz with
  * ``sha_<n>_<field>``: sha256 of the exact little-endian bytes of the state after n steps, per field
    (f, g with halo; phi with halo; convective arrays when present)
  * ``f_<N>``, ``g_<N>``, ``phi_<N>``: interior arrays after the last step N (for diagnostics and
    tolerance comparisons)
  * ``walls``: the wall file content, ``control``: the control file text
  * monitor lines (8 significant digits) from a run with the monitor switched on.
The reference is read from /root/reference only through the built binary; this script is never run on
a machine without it.
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

REF_BIN = os.path.join(REPO, "oracle", "_ref", "MF_LBM_multiphase.cpu")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


CASES = {
    # z-periodic tube + sphere, viscosity ratio 1.5, 60 degree contact angle, body-force driven slug
    "periodic_tube": dict(
        geom=lambda: geometry.tube_sphere(14, 14, 24, 5.5, 3.0),
        control={"lattice_dimensions": (14, 14, 24), "periodic_indicator": (0, 0, 1),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 4),
                 "fluid1_viscosity": 0.1, "fluid2_viscosity": 1.0 / 15.0, "surface_tension": 0.01,
                 "theta": 60.0, "body_force_0": 1e-5, "initial_fluid_distribution_option": 1,
                 "initial_interface_position": 8.0},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # velocity inlet + convective outlet through a small sphere pack, 150 degrees (drainage as in the
    # 256^3 benchmark)
    "open_velocity": dict(
        geom=lambda: geometry.sphere_pack(12, 12, 28, 0.45, 2.0, 3.5, 7, buffer=4),
        control={"lattice_dimensions": (12, 12, 28), "periodic_indicator": (0, 0, 0),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 0),
                 "inlet_BC": 1, "outlet_BC": 1, "fluid1_viscosity": 1.0 / 15.0,
                 "fluid2_viscosity": 1.0 / 15.0, "surface_tension": 0.01, "theta": 150.0,
                 "capillary_number": 0.05, "initial_fluid_distribution_option": 1,
                 "initial_interface_position": 5.0, "saturation_injection": 1.0,
                 "excluded_layers": (4, 4)},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # Zou-He pressure inlet and outlet, imbibition, 45 degrees
    "open_pressure": dict(
        geom=lambda: geometry.tube_sphere(12, 12, 24, 4.8, 2.6, buffer=3),
        control={"lattice_dimensions": (12, 12, 24), "periodic_indicator": (0, 0, 0),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 0),
                 "inlet_BC": 2, "outlet_BC": 2, "fluid1_viscosity": 0.05,
                 "fluid2_viscosity": 0.1, "surface_tension": 0.02, "theta": 45.0,
                 "body_force_0": 2e-4, "initial_fluid_distribution_option": 2,
                 "initial_interface_position": 6.0, "saturation_injection": 0.0,
                 "excluded_layers": (2, 2)},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # porous plate blocking fluid 1 near the outlet of a velocity-driven drainage (plate at z = 20)
    "porous_plate": dict(
        geom=lambda: geometry.tube_sphere(12, 12, 24, 4.8, 2.4, buffer=3),
        control={"lattice_dimensions": (12, 12, 24), "periodic_indicator": (0, 0, 0),
                 "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 0),
                 "inlet_BC": 1, "outlet_BC": 1, "fluid1_viscosity": 0.1,
                 "fluid2_viscosity": 0.1, "surface_tension": 0.01, "theta": 120.0,
                 "capillary_number": 0.05, "initial_fluid_distribution_option": 1,
                 "initial_interface_position": 14.0, "saturation_injection": 1.0,
                 "porous_plate_cmd": 1, "Z_porous_plate": 20, "excluded_layers": (3, 3)},
        steps=[0, 1, 2, 3, 40], monitor=10),
    # y- and z-periodic wall drop (contact-angle case), exercises the y faces and x-parallel edges
    "periodic_yz_drop": dict(
        geom=lambda: np.zeros((16, 14, 12), dtype=np.int8),
        control={"lattice_dimensions": (12, 14, 16), "periodic_indicator": (0, 1, 1),
                 "domain_wall_status_y": (0, 0), "domain_wall_status_z": (0, 0),
                 "MPI_async_layers_num": (0, 4, 4), "fluid1_viscosity": 1.0 / 6.0,
                 "fluid2_viscosity": 1.0 / 15.0, "surface_tension": 0.015, "theta": 70.0,
                 "body_force_0": 0.0, "initial_fluid_distribution_option": 3,
                 "initial_interface_position": 5.0},
        steps=[0, 1, 2, 3, 40], monitor=10),
}

# The 48x48x128 tube + sphere workload (tube radius 22, sphere radius 12, z-periodic, slug in the first
# 20 layers, body force 1e-6, tau = 0.7 for both fluids, gamma 0.01, 90 degrees, 2000 steps).
C1 = dict(
    geom=lambda: geometry.tube_sphere(48, 48, 128, 22.0, 12.0),
    control={"lattice_dimensions": (48, 48, 128), "periodic_indicator": (0, 0, 1),
             "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 4),
             "fluid1_viscosity": 1.0 / 15.0, "fluid2_viscosity": 1.0 / 15.0,
             "surface_tension": 0.01, "theta": 90.0, "RK_beta": 0.95, "body_force_0": 1e-6,
             "initial_fluid_distribution_option": 1, "initial_interface_position": 20.0},
    steps=[0, 2000], monitor=None, store_full=False)


def run_ref(workdir: str, walls: np.ndarray, control: dict, nsteps: int, threads: int,
            monitor: int | None = None) -> float:
    os.makedirs(workdir, exist_ok=True)
    io.write_walls(os.path.join(workdir, "walls.dat"), walls)
    ctl = dict(control)
    ctl["max_time_step"] = nsteps
    if monitor:
        ctl["monitor_timer"] = monitor
    io.write_control(os.path.join(workdir, "simulation_control.txt"), ctl)
    with open(os.path.join(workdir, "job_status.txt"), "w") as fh:
        fh.write("new_simulation\n")
    with open(os.path.join(workdir, "path_info.txt"), "w") as fh:
        fh.write("# executable location\n" + REF_BIN + "\n# geometry file\n"
                 + os.path.join(workdir, "walls.dat") + "\n# boundary file\nnone\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.time()
    r = subprocess.run([REF_BIN], cwd=workdir, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    dt = time.time() - t0
    if r.returncode != 0 or not os.path.exists(os.path.join(workdir, "out2.checkpoint", "id0000")):
        raise RuntimeError(f"reference run failed in {workdir}:\n{r.stdout[-3000:]}")
    return dt


def harvest(name: str, spec: dict, threads: int, keep: str | None) -> dict:
    walls = spec["geom"]()
    control = spec["control"]
    nx, ny, nz = control["lattice_dimensions"]
    conv = control.get("outlet_BC", 1) == 1  # allocated and saved whenever outlet_BC == 1
    out: dict = {"walls": walls, "control": json.dumps(control, sort_keys=True)}
    root = keep or tempfile.mkdtemp(prefix=f"mflbm_fix_{name}_")
    timings = {}
    for n in spec["steps"]:
        wd = os.path.join(root, f"step{n}")
        timings[n] = run_ref(wd, walls, control, n, threads)
        ck = io.read_checkpoint(os.path.join(wd, "out2.checkpoint", "id0000"), nx, ny, nz, conv)
        if ck.ntime_next != n + 1:
            raise RuntimeError(f"{name}: checkpoint after {n} steps says next step {ck.ntime_next}")
        out[f"sha_{n}_f"] = sha(ck.f)
        out[f"sha_{n}_g"] = sha(ck.g)
        out[f"sha_{n}_phi"] = sha(ck.phi)
        out[f"sha_{n}_f_int"] = sha(ck.f[:, 1:-1, 1:-1, 1:-1])
        out[f"sha_{n}_g_int"] = sha(ck.g[:, 1:-1, 1:-1, 1:-1])
        out[f"sha_{n}_phi_int"] = sha(ck.phi[4:-4, 4:-4, 4:-4])
        if conv:
            out[f"sha_{n}_convec"] = sha(np.concatenate([ck.f_convec.ravel(), ck.g_convec.ravel(),
                                                         ck.phi_convec.ravel()]))
        if n == spec["steps"][-1] and not spec.get("store_full", True):
            out[f"phi32_{n}"] = ck.phi[4:-4, 4:-4, 4:-4].astype(np.float32)
        elif n == spec["steps"][-1]:
            out[f"f_{n}"] = ck.f[:, 1:-1, 1:-1, 1:-1].copy()
            out[f"g_{n}"] = ck.g[:, 1:-1, 1:-1, 1:-1].copy()
            out[f"phi_{n}"] = ck.phi[4:-4, 4:-4, 4:-4].copy()
    if spec.get("monitor"):
        m = spec["monitor"]
        n = spec["steps"][-1]
        wd = os.path.join(root, "monitor")
        run_ref(wd, walls, control, n, threads, monitor=m)
        for fname in ("saturation.dat", "Ca_number.dat", "flowrate_time.dat", "saturation_full_domain.dat"):
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
    ap.add_argument("--c1", action="store_true", help="also run the 48x48x128 2000-step case")
    args = ap.parse_args()
    if not os.path.exists(REF_BIN):
        sys.exit("reference binary missing: run oracle/build_ref.sh first")
    todo = dict(CASES)
    if args.cases:
        todo = {k: v for k, v in CASES.items() if k in args.cases}
    if args.c1:
        todo["c1_tube_sphere_48x48x128"] = C1
    for name, spec in todo.items():
        t = time.time()
        keep = os.path.join(args.keep, name) if args.keep else None
        data = harvest(name, spec, args.threads, keep)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
        print(f"{name}: {time.time() - t:.1f}s", flush=True)


if __name__ == "__main__":
    main()
