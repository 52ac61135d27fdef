"""The run driver (mflbm/run.py) against whole runs of the reference program (tests/golden/run_variants.npz,
made by tests/golden/make_run_fixtures.py): every monitor file line for line, the stop step, the final job
status and the final checkpoint.

Here the driver steps the CPU oracle, which pins the monitor arithmetic (post_monitor: excluded layers, A_xy,
Ca, the full-domain sums, the capillary-pressure means), the stop rules (saturation, capillary-pressure and
phase-field steady states, breakthrough, failure, wall-clock stop and continuation) and the timers derived
from injected volumes against the reference's own output.  tests/test_run_driver_gpu.py runs the same
variants on the GPU path.
"""
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, sha
from mflbm import io
from mflbm.run import RunError, run_directory
from oracle_backend import OracleSimulation

RUNS = np.load(os.path.join(GOLDEN, "run_variants.npz"))
VARIANTS = sorted({k.split("/")[0] for k in RUNS.files})


def prepare(run_dir: str, name: str) -> dict:
    ctl = json.loads(str(RUNS[f"{name}/control"]))
    ctl = {k: tuple(v) if isinstance(v, list) else v for k, v in ctl.items()}
    os.makedirs(run_dir, exist_ok=True)
    io.write_walls(os.path.join(run_dir, "walls.dat"), RUNS[f"{name}/walls"])
    io.write_control(os.path.join(run_dir, "simulation_control.txt"), ctl)
    with open(os.path.join(run_dir, "job_status.txt"), "w") as fh:
        fh.write("new_simulation\n")
    with open(os.path.join(run_dir, "path_info.txt"), "w") as fh:
        fh.write("# executable location\nnone\n# geometry file\nwalls.dat\n# boundary file\nnone\n")
    return ctl


def _fields_close(a: str, b: str) -> bool:
    """Two monitor lines whose printed values agree to the last digit (a rounding flip of the 8th digit) or are
    both at round-off level (below 1e-13, where the reference's own output depends on its summation order)."""
    fa, fb = a.split(), b.split()
    if len(fa) != len(fb):
        return False
    for x, y in zip(fa, fb):
        if x == y:
            continue
        try:
            u, v = float(x), float(y)
        except ValueError:
            return False
        m = max(abs(u), abs(v))
        if not (abs(u - v) <= 2e-7 * m or m < 1e-13):
            return False
    return True


def check_against_reference(run_dir: str, name: str, ctl: dict, exact: bool = True) -> None:
    """exact=False accepts the differences a different summation order of the layer sums can make in the printed
    digits (the GPU reduces each layer in a tree); the stop step, job status and checkpoint stay exact."""
    ref_files = {k[len(name) + 6:]: str(RUNS[k]) for k in RUNS.files if k.startswith(name + "/out1/")}
    out1 = os.path.join(run_dir, "out1.output")
    ours = set()
    for root, _, files in os.walk(out1):
        for fn in files:
            if fn != "time.dat":
                ours.add(os.path.relpath(os.path.join(root, fn), out1))
    assert ours == set(ref_files), f"{name}: files written differ: {sorted(ours ^ set(ref_files))}"
    for rel, text in ref_files.items():
        with open(os.path.join(out1, rel)) as fh:
            got = fh.read().splitlines()
        want = text.splitlines()
        assert len(got) == len(want), f"{name}/{rel}: {len(got)} lines, reference {len(want)}"
        for a, b in zip(got, want):
            assert a == b or (not exact and _fields_close(a, b)), f"{name}/{rel}:\n ours      {a!r}\n reference {b!r}"
    with open(os.path.join(run_dir, "job_status.txt")) as fh:
        assert fh.read().strip() == str(RUNS[f"{name}/job_status"])
    if f"{name}/ck_next" in RUNS.files:
        nx, ny, nz = ctl["lattice_dimensions"]
        ck = io.read_checkpoint(os.path.join(run_dir, "out2.checkpoint", "id0000"), nx, ny, nz,
                                ctl.get("outlet_BC", 1) == 1)
        assert ck.ntime_next == int(RUNS[f"{name}/ck_next"])
        assert sha(ck.f) == str(RUNS[f"{name}/ck_sha_f"])
        assert sha(ck.g) == str(RUNS[f"{name}/ck_sha_g"])
        assert sha(ck.phi) == str(RUNS[f"{name}/ck_sha_phi"])


@pytest.mark.parametrize("name", VARIANTS)
def test_oracle_run_matches_reference_run(name, tmp_path):
    run_dir = str(tmp_path / name)
    ctl = prepare(run_dir, name)
    for _ in range(int(RUNS[f"{name}/invocations"])):
        run_directory(run_dir, backend=OracleSimulation)
    check_against_reference(run_dir, name, ctl)


def test_finished_runs_are_refused(tmp_path):
    """A finished, failed or unknown job status ends the program before any step (IO_multiphase.F90:27-54)."""
    run_dir = str(tmp_path / "r")
    prepare(run_dir, "tube_failure")
    for status in ("simulation_done", "simulation_reached_max_step", "simulation_failed", "bogus"):
        with open(os.path.join(run_dir, "job_status.txt"), "w") as fh:
            fh.write(status + "\n")
        with pytest.raises(RunError):
            run_directory(run_dir, backend=OracleSimulation)
    os.remove(os.path.join(run_dir, "job_status.txt"))
    with pytest.raises(RunError, match="Missing job status"):
        run_directory(run_dir, backend=OracleSimulation)
    with open(os.path.join(run_dir, "job_status.txt"), "w") as fh:
        fh.write("continue_simulation\n")
    with pytest.raises(RunError, match="Checkpoint data not found"):
        run_directory(run_dir, backend=OracleSimulation)


@pytest.mark.parametrize("name", ["periodic_tube", "open_velocity", "open_pressure", "porous_plate", "periodic_yz_drop"])
def test_oracle_monitor_files_match_fixture_runs(name, tmp_path):
    """The monitor files of the step fixtures' 40-step runs with the monitor every 10 steps
    (tests/golden/make_fixtures.py), through the run driver."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    ctl = json.loads(str(z["control"]))
    ctl = {k: tuple(v) if isinstance(v, list) else v for k, v in ctl.items()}
    ctl.update({"max_time_step": 40, "monitor_timer": 10})
    run_dir = str(tmp_path / name)
    os.makedirs(run_dir)
    io.write_walls(os.path.join(run_dir, "walls.dat"), z["walls"])
    io.write_control(os.path.join(run_dir, "simulation_control.txt"), ctl)
    with open(os.path.join(run_dir, "job_status.txt"), "w") as fh:
        fh.write("new_simulation\n")
    with open(os.path.join(run_dir, "path_info.txt"), "w") as fh:
        fh.write("#\nnone\n#\nwalls.dat\n#\nnone\n")
    r = run_directory(run_dir, backend=OracleSimulation)
    assert r.job_status == "simulation_reached_max_step" and r.ntime == 40
    for key in ("saturation", "Ca_number", "flowrate_time", "saturation_full_domain"):
        with open(os.path.join(run_dir, "out1.output", key + ".dat")) as fh:
            assert fh.read() == str(z["monitor_" + key]), key
