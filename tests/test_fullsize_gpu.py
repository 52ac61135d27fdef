"""Parity at the full benchmark size: the 256^3 sphere pack (4.4 M fluid nodes, porosity 0.22, 8 buffer layers)
stepped on the GPU and by the oracle from the same initial state, compared bit for bit.

The small fixtures pin the step against the reference's own checkpoints; this run checks that nothing in the
device path depends on the lattice size (32-bit slot offsets, tile order, the XCD block mapping, list
lengths): three steps cover an odd, an even and a second odd update with the colour-gradient chain between
them, and the monitor sums agree within 1e-12.
"""
import os

import numpy as np
import pytest

from mflbm import Case, Simulation, geometry
from mflbm.singlephase import SPCase, SPSimulation
from oracle_backend import OracleSimulation, SPOracleSimulation, set_threads

pytestmark = pytest.mark.gpu

N = 256


def _walls():
    return geometry.sphere_pack(N, N, N, 0.22, 6.0, 12.0, 2024, buffer=8)


def _threads():
    return set_threads(min(os.cpu_count() or 1, 16))


def test_two_phase_drainage_256_cube_matches_the_oracle_bit_for_bit():
    ctl = {"lattice_dimensions": (N, N, N), "periodic_indicator": (0, 0, 0), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 0), "inlet_BC": 1, "outlet_BC": 1, "fluid1_viscosity": 0.05,
           "fluid2_viscosity": 0.05, "surface_tension": 0.05, "theta": 30.0, "capillary_number": 1e-4,
           "initial_fluid_distribution_option": 1, "initial_interface_position": 8.0,
           "saturation_injection": 1.0, "excluded_layers": (8, 8)}
    case = Case(ctl, _walls())
    print(f"[full size] two-phase {N}^3, {case.n_fluid} fluid nodes, oracle threads {_threads()}", flush=True)
    st = case.initial_state()
    gpu = Simulation(case, state=st)
    ora = OracleSimulation(case, state=st)
    del st
    for sim in (gpu, ora):
        sim.run(1, 3)
    a = gpu.state()
    b = ora.state()
    for k in ("phi",):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for q in range(19):
        assert np.array_equal(a.f[q], b.f[q]), f"f{q}"
        assert np.array_equal(a.g[q], b.g[q]), f"g{q}"
    del a, b
    ma, mb = gpu.monitor(), ora.monitor()
    for key in ("saturation", "vol1_sum", "vol2_sum", "mass1_sum", "mass2_sum", "umax"):
        assert abs(getattr(ma, key) - getattr(mb, key)) <= 1e-12 * abs(getattr(mb, key)), key
    gpu.close()
    ora.close()


def test_single_phase_256_cube_matches_the_oracle_bit_for_bit():
    ctl = {"lattice_dimensions": (N, N, N), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-6,
           "MRT_collision_parameter_preset": 1}
    case = SPCase(ctl, _walls())
    print(f"[full size] single-phase {N}^3, oracle threads {_threads()}", flush=True)
    st = case.initial_state()
    gpu = SPSimulation(case, st)
    ora = SPOracleSimulation(case, st)
    del st
    for sim in (gpu, ora):
        sim.run(1, 3)
    a, b = gpu.state(), ora.state()
    for q in range(19):
        assert np.array_equal(a.f[q], b.f[q]), f"f{q}"
    ma, mb = gpu.monitor(), ora.monitor()
    assert abs(ma.flowrate - mb.flowrate) <= 1e-12 * abs(mb.flowrate)
    gpu.close()
    ora.close()
