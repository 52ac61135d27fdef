"""The algebra behind the single-phase FP32 storage mode, checked on the FP64 oracle.

The FP32 mode stores d = f - R, with R the rest equilibrium (rho = 1, u = 0) of the reference's MRT model,
and runs the unchanged collision and streaming on d.  That is exact when a time step commutes with the
shift by R: step(R + d) = R + step(d).  Both facts are checked here on the oracle, which is pinned to the
reference's checkpoints elsewhere: R is the state a uniform fluid at rest relaxes to (it is not the lattice
weights, because the reference's e2 equilibrium has no density term), and stepping R + d and d gives
results that differ by R to rounding.
"""
import numpy as np
import pytest

from mflbm.singlephase import SPCase, SPState
from oracle_backend import SPOracleSimulation

# rest equilibrium (4/21, 13/126, 1/63), the values the device header derives from the MRT transform
REST = np.array([4 / 21] + [13 / 126] * 6 + [1 / 63] * 12)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def _channel(nx, force, preset=1):
    ctl = {"lattice_dimensions": (nx, 5, 6), "periodic_indicator": (0, 1, 1), "domain_wall_status_y": (0, 0),
           "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 5, 6), "fluid_viscosity": 0.1,
           "body_force_0": force, "MRT_collision_parameter_preset": preset}
    walls = np.zeros((6, 5, nx), np.int8)
    walls[2, 2, nx // 2] = 1   # an obstacle, so the wall links of the update are exercised
    return SPCase(ctl, walls)


def _uniform(case, w):
    rp = case.rp
    st = SPState.empty(rp.nx, rp.ny, rp.nz)
    st.f[:] = w[:, None, None, None]
    return st


@pytest.mark.parametrize("preset", [1, 0])
def test_fluid_at_rest_relaxes_to_the_rest_equilibrium_not_the_lattice_weights(preset):
    case = _channel(7, 0.0, preset)
    sim = SPOracleSimulation(case, _uniform(case, W))
    sim.run(1, 1)
    f1 = sim.state().f[:, 3, 3, 2]
    assert np.abs(f1 - W).max() > 1e-3          # the lattice weights are not a fixed point
    sim.run(2, 400)
    f = sim.state().f[:, 1:-1, 1:-1, 2:-2]      # fluid nodes off the x walls
    fl = case.walls_global[:, :, 1:-1] == 0
    rho = f.sum(axis=0)[fl]                     # the wall links hold some of the mass: rho is not exactly 1
    assert np.abs(rho - 1.0).max() < 1e-2
    for q in range(19):
        np.testing.assert_allclose(f[q][fl], REST[q] * rho, rtol=0, atol=1e-15)
    assert abs(REST.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("preset", [1, 0])
def test_time_step_commutes_with_the_rest_shift(preset):
    case = _channel(8, 1e-5, preset)
    rp = case.rp
    rng = np.random.default_rng(17)
    d = 1e-3 * rng.standard_normal((19, rp.nz + 2, rp.ny + 2, rp.nx + 2))
    a = SPState(REST[:, None, None, None] + d, np.zeros((19, rp.ny + 2, rp.nx + 2)))
    b = SPState(d.copy(), np.zeros((19, rp.ny + 2, rp.nx + 2)))
    sa, sb = SPOracleSimulation(case, a), SPOracleSimulation(case, b)
    for s in (sa, sb):
        s.run(1, 6)
    fa = sa.state().f - REST[:, None, None, None]
    fb = sb.state().f
    assert np.abs(fb).max() > 1e-4
    np.testing.assert_allclose(fa, fb, rtol=0, atol=1e-15)
    # the shift by the lattice weights does not commute: the check above would catch a wrong R
    c = SPOracleSimulation(case, SPState(W[:, None, None, None] + d, np.zeros((19, rp.ny + 2, rp.nx + 2))))
    c.run(1, 6)
    assert np.abs(c.state().f - W[:, None, None, None] - fb).max() > 1e-3


def test_collision_equilibrium_state_starts_a_walled_pack_without_a_transient():
    """SPCase.collision_equilibrium_state: a fluid at rest among solid walls stays at rest (the reference's
    lattice-weight start launches density waves at every wall link), and a prescribed (rho, u) is carried by
    the populations' moments."""
    from mflbm import singlephase
    assert np.allclose(singlephase.REST, REST, rtol=0, atol=1e-16)
    assert np.allclose(singlephase.LATTICE_W, W, rtol=0, atol=1e-16)
    n = 12
    walls = (np.random.default_rng(8).random((n, n, n)) < 0.3).astype(np.int8)   # rough pore walls
    ctl = {"lattice_dimensions": (n, n, n), "periodic_indicator": (0, 0, 1), "domain_wall_status_y": (1, 1),
           "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 0, 4), "fluid_viscosity": 0.1,
           "body_force_0": 0.0, "MRT_collision_parameter_preset": 1}
    case = SPCase(ctl, walls)
    umax = {}
    for name, st in (("reference", case.initial_state()), ("equilibrium", case.collision_equilibrium_state())):
        sim = SPOracleSimulation(case, st)
        sim.run(1, 40)
        m = sim.macro()
        umax[name] = max(np.abs(m[c]).max() for c in ("u", "v", "w"))
        sim.close()
    assert umax["reference"] > 1e-3
    assert umax["equilibrium"] < 1e-14
    # moments of a prescribed state: density and momentum as given (incompressible form, u = sum e f)
    rng = np.random.default_rng(3)
    sh = (n + 2, n + 2, n + 2)
    rho = 1.0 + 1e-4 * rng.standard_normal(sh)
    vel = [1e-4 * rng.standard_normal(sh) for _ in range(3)]
    f = np.array(case.collision_equilibrium_state(rho=rho, velocity=vel).f)
    ex = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
    ey = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
    ez = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
    opp = [0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15]
    fl = np.zeros(f.shape[1:], bool)
    fl[1:-1, 1:-1, 1:-1] = case.walls_global == 0
    np.testing.assert_allclose(f.sum(axis=0)[fl], rho[fl], rtol=0, atol=1e-14)
    for e, v in zip((ex, ey, ez), vel):
        np.testing.assert_allclose(np.tensordot(e, f, 1)[fl], (rho * v)[fl], rtol=0, atol=1e-14)
    # a solid slot facing a fluid node holds that node's opposite population (the first step's bounce-back)
    rng2 = np.random.default_rng(5)
    solid = np.argwhere(~fl[1:-1, 1:-1, 1:-1]) + 1
    checked = 0
    for z, y, x in solid[rng2.permutation(len(solid))[:200]]:
        for q in range(1, 19):
            zz, yy, xx = z + ez[q], y + ey[q], x + ex[q]
            if fl[zz, yy, xx]:
                assert f[q, z, y, x] == f[opp[q], zz, yy, xx]
                checked += 1
    assert checked > 100
