"""The colour-gradient chain skips bricks (16 x 2 x 2 nodes) whose neighbourhood holds phi within 3e-7 of one of
+-1 (there every normal is +0), leaves the extrapolated phi of solid nodes there as it was until a download needs
it, and the lattice update skips the normal stencil in bricks with no set bit nearby.  These tests run
domains large enough for most bricks to be skipped, with monitor calls in between and at the last step (the
macroscopic pass zeroes phi at solid nodes), and check:

* FP64: every population and phi bit for bit against the oracle, which evaluates every node every step;
* FP32 storage (FP32 and mixed collision, brick plan forced with MFLBM_CHAIN=bricks): bit for bit against the
  same run with a full chain forced after every step (mflbm_color_gradient evaluates every brick), so a skipped
  brick that should have been evaluated shows as a difference, and against the chain over whole lists
  (MFLBM_CHAIN=all, the FP32-collision default), which evaluates every entry every step.
"""
import numpy as np
import pytest

from mflbm import Case, Simulation, geometry
from oracle_backend import OracleSimulation, set_threads

pytestmark = pytest.mark.gpu

MONITOR_AT = (25, 60, 61, 140, 160, 200)


def _drainage(precision="fp64"):
    n = (40, 40, 96)
    walls = geometry.sphere_pack(*n, 0.35, 3.0, 6.0, 31, buffer=6)
    ctl = {"lattice_dimensions": n, "periodic_indicator": (0, 0, 0), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 0), "inlet_BC": 1, "outlet_BC": 1, "fluid1_viscosity": 0.1,
           "fluid2_viscosity": 0.1, "surface_tension": 0.02, "theta": 40.0, "capillary_number": 2e-3,
           "initial_fluid_distribution_option": 1, "initial_interface_position": 6.0, "saturation_injection": 1.0,
           "excluded_layers": (6, 6)}
    return Case(ctl, walls, precision=precision)


def _periodic_slug(precision="fp64"):
    ctl = {"lattice_dimensions": (48, 48, 128), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 4), "fluid1_viscosity": 1.0 / 15.0, "fluid2_viscosity": 0.1,
           "surface_tension": 0.01, "theta": 60.0, "body_force_0": 2e-5, "initial_fluid_distribution_option": 1,
           "initial_interface_position": 20.0}
    return Case(ctl, geometry.tube_sphere(48, 48, 128, 22.0, 12.0), precision=precision)


def _run(sim, nsteps, full_every_step=False):
    for n in range(1, nsteps + 1):
        sim.step(n)
        if full_every_step:
            sim.color_gradient()
        if n in MONITOR_AT:
            sim.monitor(update_phi_old=True)


def _same(a, b, what):
    for key in ("f", "g", "phi", "f_convec", "g_convec", "phi_convec"):
        x, y = getattr(a, key), getattr(b, key)
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), f"{what}: {key} differs"


@pytest.mark.parametrize("make,nsteps", [(_drainage, 200), (_periodic_slug, 160)])
def test_brick_skipping_chain_is_bit_identical_to_the_oracle(make, nsteps):
    case = make()
    print(f"[bricks fp64] {make.__name__} {nsteps} steps", flush=True)
    set_threads(16)
    gpu, ora = Simulation(case), OracleSimulation(case)
    _run(gpu, nsteps)
    _run(ora, nsteps)
    _same(gpu.state(), ora.state(), make.__name__)
    gg, go = gpu.gradient(), ora.gradient()
    for k in gg:
        assert np.array_equal(gg[k], go[k]), k
    mg, mo = gpu.macro(), ora.macro()
    for k in ("u", "v", "w", "rho"):
        assert np.array_equal(mg[k][1:-1, 1:-1, 1:-1], mo[k][1:-1, 1:-1, 1:-1]), k
    assert np.array_equal(mg["phi"].view(np.uint64), mo["phi"].view(np.uint64))
    gpu.close()
    ora.close()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("make,nsteps", [(_drainage, 200), (_periodic_slug, 160)])
def test_brick_skipping_chain_equals_a_full_chain_with_fp32_storage(make, nsteps, precision, monkeypatch):
    print(f"[bricks {precision}] {make.__name__} {nsteps} steps", flush=True)
    monkeypatch.setenv("MFLBM_CHAIN", "bricks")
    a, b = Simulation(make(precision)), Simulation(make(precision))
    monkeypatch.setenv("MFLBM_CHAIN", "all")
    w = Simulation(make(precision))
    monkeypatch.delenv("MFLBM_CHAIN")
    _run(a, nsteps)
    _run(b, nsteps, full_every_step=True)
    _run(w, nsteps)
    sa = a.state()
    _same(sa, b.state(), f"{make.__name__} {precision}")
    _same(sa, w.state(), f"{make.__name__} {precision} whole lists")
    for s in (a, b, w):
        s.close()
