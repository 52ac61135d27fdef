"""gfx950 path against the reference's checkpoints and against the oracle (bit-for-bit, FP64).

The step is FP64 throughout and evaluates the reference's expressions in the reference's order without
FMA contraction, so the expected agreement is exact equality of every population bit; the tolerance
named by the task (1e-12 relative on phi, velocity and saturation) is checked as well where a reduction
order differs (monitor sums).
"""
import numpy as np
import pytest

from helpers import (C1_CASE, SMALL_CASES, assert_matches_reference, fixture_case, fixture_steps, max_rel_diff,
                     phi_as_reference_keeps_it, sha)
from mflbm import Simulation
from oracle_backend import OracleSimulation, set_threads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", SMALL_CASES)
def test_gpu_trajectory_matches_reference(name):
    z, case = fixture_case(name)
    print(f"[gpu parity] case={name}", flush=True)
    sim = Simulation(case)
    t = 0
    for n in fixture_steps(z):
        if n > t:
            sim.run(t + 1, n - t)
            t = n
        assert_matches_reference(z, case, sim.state(), n)
    sim.close()


@pytest.mark.parametrize("name", SMALL_CASES)
def test_gpu_gradient_and_macro_match_oracle(name):
    z, case = fixture_case(name)
    print(f"[gpu vs oracle fields] case={name}", flush=True)
    gpu, ora = Simulation(case), OracleSimulation(case)
    for sim in (gpu, ora):
        sim.run(1, 24)
    gg, go = gpu.gradient(), ora.gradient()
    for k in gg:
        assert np.array_equal(gg[k], go[k]), k
    mg, mo = gpu.macro(), ora.macro()
    for k in ("u", "v", "w", "rho"):
        assert np.array_equal(mg[k][1:-1, 1:-1, 1:-1], mo[k][1:-1, 1:-1, 1:-1]), k
    gpu.close()


@pytest.mark.parametrize("name", SMALL_CASES)
def test_gpu_monitor_matches_oracle(name):
    z, case = fixture_case(name)
    gpu, ora = Simulation(case), OracleSimulation(case)
    for sim in (gpu, ora):
        sim.run(1, 10)
        sim.monitor(update_phi_old=True)
        sim.run(11, 10)
    a, b = gpu.monitor(), ora.monitor()
    for key in ("saturation", "saturation_full_domain", "vol1_sum", "vol2_sum", "mass1_sum", "mass2_sum"):
        assert abs(getattr(a, key) - getattr(b, key)) <= 1e-12 * abs(getattr(b, key)), key
    # Ca is a signed sum of fluxes; where the net flux is round-off (a drop at rest) the tolerance is taken
    # relative to the sum of the magnitudes of the layer fluxes it is formed from
    rp = case.rp
    flux_scale = (np.abs(b.layers[0]).sum() + np.abs(b.layers[1]).sum()) / rp.nz / rp.A_xy * rp.la_nu1 / rp.gamma
    assert abs(a.ca - b.ca) <= 1e-12 * max(abs(b.ca), flux_scale)
    assert a.umax == b.umax
    assert a.dphi_max == b.dphi_max
    assert a.outlet_phase1 == b.outlet_phase1
    for m in range(7):
        np.testing.assert_allclose(a.layers[m], b.layers[m], rtol=1e-12, atol=1e-12 * np.abs(b.layers[m]).max())
    gpu.close()


def test_gpu_c1_2000_steps_matches_reference_and_oracle():
    z, case = fixture_case(C1_CASE)
    print("[gpu parity] 48x48x128 tube+sphere, 2000 steps", flush=True)
    sim = Simulation(case)
    sim.run(1, 2000)
    st = sim.state()
    assert_matches_reference(z, case, st, 2000)
    # saturation and velocity against the oracle's monitor at the same state
    set_threads(8)
    ora = OracleSimulation(case, state=st)
    a, b = sim.monitor(), ora.monitor()
    assert abs(a.saturation_full_domain - b.saturation_full_domain) <= 1e-12 * b.saturation_full_domain
    mg, mo = sim.macro(), ora.macro()
    for k in ("u", "v", "w"):
        assert max_rel_diff(mg[k][1:-1, 1:-1, 1:-1], mo[k][1:-1, 1:-1, 1:-1]) <= 1e-12
    sim.close()


@pytest.mark.parametrize("name", ["open_velocity", "periodic_yz_drop"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_gradient_chain_equals_the_list_kernels(name, precision, monkeypatch):
    """MFLBM_CHAIN=fused: normals, extrapolation and curvature force in one LDS-tiled pass (k_chain_force), and
    MFLBM_CHAIN=all: the list kernels over every entry without the brick plan; the populations and phi after a run
    are bit-identical to the brick-listed kernels' and the downloaded gradient too."""
    from helpers import load_fixture
    from mflbm import Case
    z, ctl = load_fixture(name)
    case = Case(ctl, z["walls"], precision=precision)
    print(f"[fused chain] case={name} precision={precision}", flush=True)
    sims = {}
    for chain in ("bricks", "all", "fused"):
        monkeypatch.setenv("MFLBM_CHAIN", chain)
        sim = Simulation(case)
        sim.run(1, 12)
        sims[chain] = (sim.state(), sim.gradient(), sim.monitor())
        sim.close()
    a = sims["bricks"]
    for chain in ("all", "fused"):
        b = sims[chain]
        for k in ("f", "g", "phi"):
            assert np.array_equal(getattr(a[0], k), getattr(b[0], k)), (chain, k)
        for k in a[1]:
            assert np.array_equal(a[1][k], b[1][k]), (chain, k)
        assert a[2].saturation == b[2].saturation, chain
