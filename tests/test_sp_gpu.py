"""gfx950 single-phase path (mflbm_sp_*) against the reference's checkpoints and the oracle.

FP64 throughout, the reference's expressions in the reference's order without FMA contraction: the expected
agreement is equality of every population bit.  The monitor's layer sums are tree reductions on the device
(the reference adds serially), so they are checked to 1e-12 relative of the summed magnitudes.
"""
import numpy as np
import pytest

from helpers import SP_CASES, assert_sp_matches, fixture_steps, load_sp, sha
from mflbm import geometry
from mflbm.singlephase import SPCase, SPSimulation
from oracle_backend import SPOracleSimulation, set_threads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", SP_CASES)
def test_sp_gpu_trajectory_matches_reference(name):
    z, case = load_sp(name)
    print(f"[sp gpu parity] case={name}", flush=True)
    sim = SPSimulation(case)
    t = 0
    for n in fixture_steps(z):
        if n > t:
            sim.run(t + 1, n - t)
            t = n
        assert_sp_matches(z, case, sim.state(), n)
    sim.close()


@pytest.mark.parametrize("name", SP_CASES)
def test_sp_gpu_macro_and_monitor_match_oracle(name):
    z, case = load_sp(name)
    print(f"[sp gpu vs oracle] case={name}", flush=True)
    gpu, ora = SPSimulation(case), SPOracleSimulation(case)
    for sim in (gpu, ora):
        sim.run(1, 30)
    mg, mo = gpu.macro(), ora.macro()
    for k in ("u", "v", "w", "rho"):
        assert np.array_equal(mg[k][1:-1, 1:-1, 1:-1], mo[k][1:-1, 1:-1, 1:-1]), k
    a, b = gpu.monitor(), ora.monitor()
    assert a.umax == b.umax
    scale_fl = np.abs(mo["w"]).sum() / case.rp.nz + 1e-300
    assert abs(a.flowrate - b.flowrate) <= 1e-12 * scale_fl
    np.testing.assert_allclose(a.layers[1], b.layers[1], rtol=1e-12)
    if b.pressure is not None:
        for x, y in zip(a.pressure[:2], b.pressure[:2]):
            assert abs(x - y) <= 1e-12 * abs(y)
    gpu.close()


def test_sp_gpu_plane_poiseuille_is_exact():
    nx, nu, force = 10, 0.1, 1e-5
    ctl = {"lattice_dimensions": (nx, 4, 4), "periodic_indicator": (0, 1, 1), "domain_wall_status_y": (0, 0),
           "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 4, 4), "fluid_viscosity": nu,
           "body_force_0": force, "MRT_collision_parameter_preset": 1, "steady_state_option": 1,
           "convergence_criteria": 1e-13}
    case = SPCase(ctl, np.zeros((4, 4, nx), np.int8))
    sim = SPSimulation(case)
    n, status, hist = sim.run_steady(1, 100000, 200)
    assert status == "converged"
    x = np.arange(1, nx + 1) - 1.5
    wa = force / (2.0 * nu) * x * (nx - 2 - x)
    wa[0] = wa[-1] = 0.0
    w = sim.macro()["w"][2, 2, 1:-1]
    np.testing.assert_allclose(w, wa, rtol=0, atol=1e-11 * wa.max())


def test_sp_gpu_sphere_pack_matches_oracle_bit_for_bit():
    # 48 x 48 x 96 z-periodic sphere pack (sample porosity 0.35, 8 fluid buffer layers at each end), 120 steps
    walls = geometry.sphere_pack(48, 48, 96, 0.35, 4.0, 8.0, 11)
    ctl = {"lattice_dimensions": (48, 48, 96), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-5,
           "MRT_collision_parameter_preset": 1}
    case = SPCase(ctl, walls)
    print(f"[sp gpu pack] fluid nodes={case.n_fluid}", flush=True)
    set_threads(16)
    gpu, ora = SPSimulation(case), SPOracleSimulation(case)
    for sim in (gpu, ora):
        sim.run(1, 120)
    assert sha(gpu.state().f) == sha(ora.state().f)
    gpu.close()
