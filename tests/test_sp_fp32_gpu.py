"""Single-phase FP32 deviation storage on the GPU (mflbm_sp_params.precision = 1: collided in FP32; 2, "mixed":
collided in FP64).

The populations are kept as d = f - R in FP32 (R the rest equilibrium of the MRT model); boundaries,
macroscopic fields and reductions stay FP64.  Checks:
* a fluid at rest is kept exactly: R is stored as d = 0, which the FP32 step maps to 0, and the download adds R
  back in FP64, so any difference between the device's R and the rest equilibrium shows up bit for bit;
* trajectories of the reference's single-phase cases stay within FP32 rounding of the FP64 oracle (the
  tolerances are stated per test);
* a sphere-pack permeability run converged to the reference's criterion 1e-7 in the mixed mode gives the FP64
  path's permeability to better than 0.5 %; in the FP32 mode the flow rate keeps fluctuating by about 1e-5
  (FP32 collision rounding; see the mass drift it also shows), so that run is compared after as many steps as
  the FP64 run needed.
"""
import numpy as np
import pytest

from helpers import SP_CASES, load_sp
from mflbm import geometry
from mflbm.singlephase import SPCase, SPSimulation, SPState, permeability
from oracle_backend import SPOracleSimulation

pytestmark = pytest.mark.gpu

REST = np.array([4 / 21] + [13 / 126] * 6 + [1 / 63] * 12)


def test_sp_fp32_keeps_a_fluid_at_rest_exactly():
    ctl = {"lattice_dimensions": (9, 6, 7), "periodic_indicator": (0, 1, 1), "domain_wall_status_y": (0, 0),
           "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 6, 7), "fluid_viscosity": 0.1,
           "body_force_0": 0.0, "MRT_collision_parameter_preset": 1}
    walls = np.zeros((7, 6, 9), np.int8)
    walls[3, 2:4, 4] = 1
    case = SPCase(ctl, walls, "fp32")
    rp = case.rp
    st = SPState.empty(rp.nx, rp.ny, rp.nz)
    st.f[:] = REST[:, None, None, None]
    sim = SPSimulation(case, st)
    sim.run(1, 10)
    f = sim.state().f
    assert np.array_equal(f, st.f)
    m = sim.macro()
    fl = case.walls_global == 0
    rho = 0.0
    for q in range(19):   # the macro kernel's order: f0 + f1 + ... + f18
        rho += REST[q]
    assert np.array_equal(m["rho"][1:-1, 1:-1, 1:-1][fl], np.full(fl.sum(), rho))
    assert not np.any(m["w"])
    sim.close()


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("name", SP_CASES)
def test_sp_fp32_trajectory_tracks_the_fp64_oracle(name, prec):
    """60 steps from the reference's initial state.  Each stored deviation carries a relative error of 2^-24,
    and the first steps hold deviations up to 0.15 (the initial state is built on the lattice weights, not on
    the rest equilibrium): about 1e-8 per population per step before the e2 mode has relaxed, far less after.
    The errors scale with the flow: bounds are 5e-6 (populations, velocity) and 3e-5 (density) times the
    largest velocity component of the FP64 run (measured: at most 1.3e-6, 6e-7 and 7e-6 of it)."""
    z, c64 = load_sp(name)
    _, c32 = load_sp(name, prec)
    print(f"[sp {prec} vs oracle] case={name}", flush=True)
    g, o = SPSimulation(c32), SPOracleSimulation(c64)
    for sim in (g, o):
        sim.run(1, 60)
    fg, fo = g.state().f, o.state().f
    fl = c64.walls_global == 0
    df = np.abs(fg - fo)[:, 1:-1, 1:-1, 1:-1][:, fl].max()
    mg, mo = g.macro(), o.macro()
    uscale = max(np.abs(mo[k][1:-1, 1:-1, 1:-1][fl]).max() for k in ("u", "v", "w"))
    du = max(np.abs(mg[k] - mo[k])[1:-1, 1:-1, 1:-1][fl].max() for k in ("u", "v", "w"))
    drho = np.abs(mg["rho"] - mo["rho"])[1:-1, 1:-1, 1:-1][fl].max()
    print(f"[sp {prec} vs oracle] {name}: max|df|={df:.3e} max|du|={du:.3e} (u scale {uscale:.3e}) "
          f"max|drho|={drho:.3e}", flush=True)
    assert df < 5e-6 * uscale
    assert drho < 3e-5 * uscale
    assert du < 5e-6 * uscale
    a, b = g.monitor(), o.monitor()
    assert abs(a.flowrate - b.flowrate) <= 1e-5 * abs(b.flowrate) + 1e-9
    g.close()


def test_sp_fp32_converged_permeability_matches_fp64():
    """48 x 48 x 96 z-periodic sphere pack (porosity 0.3, 8 fluid buffer layers at each end), body force 1e-6,
    run to the reference's flow-rate criterion 1e-7 (Monitor.F90:145-156) in both precisions."""
    walls = geometry.sphere_pack(48, 48, 96, 0.3, 4.0, 8.0, 5)
    ctl = {"lattice_dimensions": (48, 48, 96), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-6,
           "MRT_collision_parameter_preset": 1, "steady_state_option": 1, "convergence_criteria": 1e-7}
    k = {}
    steps = {}
    for prec in ("fp64", "mixed"):
        case = SPCase(ctl, walls, prec)
        sim = SPSimulation(case)
        n, status, hist = sim.run_steady(1, 60000, 500)
        assert status == "converged", (prec, n, hist[-3:])
        k[prec] = permeability(case, hist[-1][1])
        steps[prec] = n
        sim.close()
    case = SPCase(ctl, walls, "fp32")
    sim = SPSimulation(case)
    sim.run(1, steps["fp64"])
    k["fp32"] = permeability(case, sim.monitor().flowrate)
    sim.close()
    rel = {p: abs(k[p] / k["fp64"] - 1.0) for p in ("mixed", "fp32")}
    print(f"[sp fp32 permeability] k={k} rel={rel} steps={steps}", flush=True)
    assert rel["mixed"] < 5e-3 and rel["fp32"] < 5e-3
