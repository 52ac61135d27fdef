"""Initial state of a new run computed on the device (mflbm_init_state_device) against the reference's
step-0 checkpoints and against the host initialisation uploaded through mflbm_upload_state, bit for bit."""
import numpy as np
import pytest

from helpers import SMALL_CASES, assert_matches_reference, fixture_case
from mflbm import Simulation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", SMALL_CASES + ["c1_tube_sphere_48x48x128"])
def test_device_initial_state_equals_the_reference_and_the_host_path(name):
    z, case = fixture_case(name)
    print(f"[device init] case={name}", flush=True)
    dev = Simulation(case, device_init=True, gradient=False)
    host = Simulation(case, gradient=False)
    a, b = dev.state(), host.state()
    assert_matches_reference(z, case, a, 0)
    for k in ("f", "g", "phi", "f_convec", "g_convec", "phi_convec"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    # and the runs continue identically
    for s in (dev, host):
        s.color_gradient()
        s.run(1, 6)
    a, b = dev.state(), host.state()
    for k in ("f", "g", "phi"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    dev.close()
    host.close()
