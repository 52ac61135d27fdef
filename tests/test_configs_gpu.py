"""The three configurations that layer additions on the FP64 path, at their stated sizes.

* FP32 storage on the 256^3 drainage: FP32 and mixed (FP32 storage, FP64 collision) modes against the FP64 path
  from the same device-built initial state (saturation, capillary number, largest velocity after 1000 and 2000
  steps), and the spurious current of a static drop (64^3, radius 16, 90 degrees, 4000 steps).
* Batched REVs of sides 48, 64, 96 and 128 in one batch: every sample's populations after a few steps
  against the oracle (FP64 bit for bit; FP32 storage within its rounding).
* Learned warm start trained on 64^3 fields (a reduced count) and used on unseen 128^3 and 256^3 packs: the
  restarted runs meet the reference's 1e-7 criterion at the plain path's permeability within 0.5 %.

Bounds are stated with each test; the measured values print with the test's name.
"""
import os

import numpy as np
import pytest

from mflbm import Case, Simulation, batch, geometry
from mflbm.batch import REVBatch, REVSample, global_walls
from mflbm.singlephase import SPCase
from oracle_backend import SPOracleSimulation, set_threads

pytestmark = pytest.mark.gpu


def _drainage(precision: str) -> Case:
    n = 256
    walls = geometry.sphere_pack(n, n, n, 0.22, 6.0, 12.0, 2024, buffer=8)
    ctl = {"lattice_dimensions": (n, n, n), "periodic_indicator": (0, 0, 0), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 0), "inlet_BC": 1, "outlet_BC": 1, "fluid1_viscosity": 0.05,
           "fluid2_viscosity": 0.05, "surface_tension": 0.05, "theta": 30.0, "capillary_number": 1e-4,
           "initial_fluid_distribution_option": 1, "initial_interface_position": 8.0,
           "saturation_injection": 1.0, "excluded_layers": (8, 8)}
    return Case(ctl, walls, geometry="device", precision=precision)


def _drop(precision: str) -> Case:
    ctl = {"lattice_dimensions": (64, 64, 64), "periodic_indicator": (0, 1, 1), "domain_wall_status_y": (0, 0),
           "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 4, 4), "fluid1_viscosity": 1.0 / 6.0,
           "fluid2_viscosity": 1.0 / 6.0, "surface_tension": 0.01, "theta": 90.0, "body_force_0": 0.0,
           "initial_fluid_distribution_option": 5, "initial_interface_position": 16.0}
    return Case(ctl, np.zeros((64, 64, 64), np.int8), precision=precision)


# bounds per mode: saturation (absolute), capillary number and largest velocity (relative), spurious current
# (relative); the FP32 collision's rounding (about 6e-8 relative per operation) accumulates over the steps,
# the mixed mode only rounds the stored populations
# (measured after 2000 steps: FP32 saturation 1.3e-5, Ca 2.1e-3 while the start-up transient still dominates the
# net flux, umax 1.2e-5)
BOUNDS = {"fp32": {"sat": 1e-4, "ca": 5e-3, "umax": 2e-3, "spur": 2e-2},
          "mixed": {"sat": 2e-6, "ca": 2e-3, "umax": 2e-4, "spur": 1e-4}}


def test_fp32_storage_on_the_256_cube_drainage_tracks_fp64():
    rows = {}
    for prec in ("fp64", "fp32", "mixed"):
        print(f"[C3] 256^3 drainage, {prec}", flush=True)
        sim = Simulation(_drainage(prec), device_init=True)
        rows[prec] = []
        t = 0
        for n in (1000, 2000):
            sim.run(t + 1, n - t)
            t = n
            m = sim.monitor()
            assert not m.failed
            rows[prec].append((m.saturation, m.ca, m.umax))
        sim.close()
    bad = []
    for prec in ("fp32", "mixed"):
        b = BOUNDS[prec]
        for k, n in enumerate((1000, 2000)):
            s, ca, um = rows[prec][k]
            s0, ca0, um0 = rows["fp64"][k]
            e = (abs(s - s0), abs(ca / ca0 - 1.0), abs(um / um0 - 1.0))
            print(f"[C3] {prec} step {n}: saturation {s:.8f} (fp64 {s0:.8f}) abs err {e[0]:.2e}, Ca rel err {e[1]:.2e}, "
                  f"umax rel err {e[2]:.2e}", flush=True)
            if not (e[0] <= b["sat"] and e[1] <= b["ca"] and e[2] <= b["umax"]):
                bad.append((prec, n, e))
    assert not bad, bad


def test_fp32_storage_spurious_current_of_a_static_drop():
    spur = {}
    for prec in ("fp64", "fp32", "mixed"):
        sim = Simulation(_drop(prec), device_init=True)
        sim.run(1, 4000)
        spur[prec] = sim.monitor().umax
        sim.close()
    for prec in ("fp32", "mixed"):
        rel = abs(spur[prec] / spur["fp64"] - 1.0)
        print(f"[C3] static drop 64^3, 4000 steps: spurious current {prec} {spur[prec]:.6e} vs fp64 "
              f"{spur['fp64']:.6e} (rel {rel:.2e})", flush=True)
        assert rel <= BOUNDS[prec]["spur"]


def _rev_case(s: REVSample, precision: str) -> SPCase:
    nx, ny, nz = s.shape
    ctl = {"lattice_dimensions": (nx, ny, nz), "periodic_indicator": (0, s.jper, 1),
           "domain_wall_status_y": (0, 0) if s.jper else (1, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 4 if s.jper else 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-6,
           "MRT_collision_parameter_preset": 1, "steady_state_option": 1, "convergence_criteria": 1e-7}
    return SPCase(ctl, s.walls, precision)


def _c4_samples():
    return [batch.rev_samples(1, seed=500 + k, sides=(side,))[0] for k, side in enumerate((48, 64, 96, 128))]


def test_batch_of_sides_48_to_128_matches_the_oracle_bit_for_bit():
    """One FP64 batch holding a REV of each side of the batched configuration (48, 64, 96, 128; sphere packs of
    porosity 0.15-0.35), 6 steps from the reference's initial state, against the single-phase oracle run on
    each sample alone: every population bit for bit."""
    samples = _c4_samples()
    set_threads(min(os.cpu_count() or 1, 16))
    b = REVBatch(samples, 0.1, 1e-6, "fp64")
    b.run(1, 6)
    for i, s in enumerate(samples):
        print(f"[C4] fp64 sample {i}: {s.shape}", flush=True)
        ora = SPOracleSimulation(_rev_case(s, "fp64"))
        ora.run(1, 6)
        fo = ora.state().f
        fb = b.state(i)
        fl = global_walls(s) == 0
        assert b.sample_fluid(i) == int(fl.sum())
        for q in range(19):
            assert np.array_equal(fb[q, 1:-1, 1:-1, 1:-1][fl], fo[q, 1:-1, 1:-1, 1:-1][fl]), (i, q)
        ora.close()
    b.close()


def test_fp32_batch_of_sides_48_to_128_tracks_the_oracle():
    """The same batch with FP32 storage and FP32 collision, started at the collision's rest state, 40 steps,
    against the FP64 oracle from the same state: velocity within 1e-3 of each sample's velocity scale, density
    within 1e-6."""
    samples = _c4_samples()
    set_threads(min(os.cpu_count() or 1, 16))
    b = REVBatch(samples, 0.1, 1e-6, "fp32", rest_state=1)
    b.run(1, 40)
    for i, s in enumerate(samples):
        case = _rev_case(s, "fp64")
        ora = SPOracleSimulation(case, case.collision_equilibrium_state())
        ora.run(1, 40)
        mo, mb = ora.macro(), b.macro(i)
        fl = np.pad(global_walls(s) == 0, 1)
        us = max(np.abs(mo[k][fl]).max() for k in ("u", "v", "w"))
        du = max(np.abs(mb[k][fl] - mo[k][fl]).max() for k in ("u", "v", "w"))
        dr = np.abs(mb["rho"][fl] - mo["rho"][fl]).max()
        print(f"[C4] fp32 sample {i} {s.shape}: max|du| {du:.3e} (u scale {us:.3e}), max|drho| {dr:.3e}", flush=True)
        assert du <= 1e-3 * us and dr <= 1e-6
        ora.close()
    b.close()


def test_warm_start_trained_at_64_restarts_128_and_256_packs_to_the_criterion():
    """Trained on 64^3 fields (64 packs, 12 epochs: a reduced count of the 2000-field configuration), used on
    unseen packs of side 128 and 256: the FP32 run restarted from the network's fields meets the reference's
    1e-7 flow-rate criterion, at the plain FP64 path's permeability within 0.5 %.  Steps and wall times of the
    plain, rest-state and network starts print; the network's own gain in steps is not asserted here (see the
    warm-start gain test)."""
    from mflbm import warmstart
    data = warmstart.make_dataset(64, side=64, seed=11)
    net = warmstart.WarmStartNet(seed=0)
    hist = warmstart.train(net, data, epochs=12, batch_size=4, lr=1e-3)
    print(f"[C5] trained on {len(data.conv_steps)} fields of 64^3: loss {hist[0]:.4g} -> {hist[-1]:.4g}", flush=True)
    for k, side in enumerate((128, 256)):
        s = batch.rev_samples(1, seed=90001 + k, sides=(side,), porosity=(0.25, 0.25))[0]
        r = warmstart.compare_runs(net, s)
        print(f"[C5] {side}^3: steps plain {r.plain.steps} cold {r.cold.steps} warm {r.warm.steps}; seconds plain "
              f"{r.plain.seconds:.2f} cold {r.cold.seconds:.2f} warm {r.warm.seconds:.2f}; k rel err {r.k_rel_err:.2e}",
              flush=True)
        for rec in (r.plain, r.cold, r.warm):
            assert rec.status == "converged"
        assert r.k_rel_err < 5e-3
