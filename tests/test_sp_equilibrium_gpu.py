"""mflbm_sp_set_equilibrium: the collision-equilibrium state of prescribed fields built on the device equals the
host-built SPCase.collision_equilibrium_state uploaded through mflbm_sp_upload_state, bit for bit (populations,
wall slots included), in FP64 and in the FP32 storage modes."""
import numpy as np
import pytest

from mflbm.singlephase import SPCase, SPSimulation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
@pytest.mark.parametrize("jper", [0, 1])
def test_device_equilibrium_state_equals_the_host_built_one(precision, jper):
    nx, ny, nz = 20, 22, 24
    rng = np.random.default_rng(11 + jper)
    walls = (rng.random((nz, ny, nx)) < 0.3).astype(np.int8)
    ctl = {"lattice_dimensions": (nx, ny, nz), "periodic_indicator": (0, jper, 1),
           "domain_wall_status_y": (0, 0) if jper else (1, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 4 if jper else 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-6,
           "MRT_collision_parameter_preset": 1}
    case = SPCase(ctl, walls, precision)
    sh = (nz + 2, ny + 2, nx + 2)
    rho = 1.0 + 1e-4 * rng.standard_normal(sh)
    vel = [1e-4 * rng.standard_normal(sh) for _ in range(3)]
    print(f"[sp equilibrium] precision={precision} jper={jper}", flush=True)
    host = SPSimulation(case, case.collision_equilibrium_state(rho=rho, velocity=vel))
    dev = SPSimulation(case)
    dev.set_equilibrium(rho, vel)
    a, b = host.state().f, dev.state().f
    assert np.array_equal(a, b)
    # and both runs continue identically
    for s in (host, dev):
        s.run(1, 9)
    assert np.array_equal(host.state().f, dev.state().f)
    host.close()
    dev.close()
