"""The oracle (with the product's host set-up) against checkpoints of the unmodified reference solver.

The fixtures hold sha256 digests of the reference's exact checkpoint bytes after 0, 1, 2, 3 and 40 steps
(and 2000 steps of the 48x48x128 tube + sphere case); a match is bit-for-bit.  Covered: z-periodic
body-force drive with viscosity contrast and 60 degree wetting; velocity inlet + convective outlet through a
sphere pack at 150 degrees; Zou-He pressure inlet and outlet (imbibition, 45 degrees); y- and z-periodic
wall drop (faces and x-parallel edges of the periodic exchange).
"""
import numpy as np
import pytest

from helpers import C1_CASE, SMALL_CASES, assert_matches_reference, fixture_case, fixture_steps, sha
from oracle_backend import OracleSimulation, set_threads


@pytest.mark.parametrize("name", SMALL_CASES)
def test_initial_state_matches_reference(name):
    z, case = fixture_case(name)
    st = case.initial_state()
    assert sha(st.f) == str(z["sha_0_f"])
    assert sha(st.g) == str(z["sha_0_g"])


@pytest.mark.parametrize("name", SMALL_CASES)
def test_trajectory_matches_reference(name):
    z, case = fixture_case(name)
    sim = OracleSimulation(case)
    t = 0
    for n in fixture_steps(z):
        if n > t:
            sim.run(t + 1, n - t)
            t = n
        assert_matches_reference(z, case, sim.state(), n)
    if "convec" in " ".join(z.files) and f"sha_{t}_convec" in z.files:
        st = sim.state()
        got = sha(np.concatenate([st.f_convec.ravel(), st.g_convec.ravel(), st.phi_convec.ravel()]))
        assert got == str(z[f"sha_{t}_convec"])


def test_c1_2000_steps_matches_reference():
    z, case = fixture_case(C1_CASE)
    set_threads(8)
    sim = OracleSimulation(case)
    sim.run(1, 2000)
    assert_matches_reference(z, case, sim.state(), 2000)


def test_thread_count_does_not_change_the_state():
    z, case = fixture_case("open_velocity")
    out = []
    for nt in (1, 4):
        set_threads(nt)
        sim = OracleSimulation(case)
        sim.run(1, 12)
        st = sim.state()
        out.append((sha(st.f), sha(st.g), sha(st.phi)))
    set_threads(8)
    assert out[0] == out[1]
