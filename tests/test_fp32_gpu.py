"""Two-phase FP32 storage (mflbm_params.precision = 1) against the FP64 oracle.

The populations are kept as FP32 s = f + g - REST and a = f - g, collided and recoloured in FP32; phi, the
colour-gradient chain, boundaries and monitors stay FP64.  The reference's small cases are run for 40 steps
from the same initial state; phi, velocity and density must stay within FP32 rounding of the FP64 run (bounds
stated in the test), and the saturation to 1e-6.
"""
import numpy as np
import pytest

from helpers import SMALL_CASES, load_fixture
from mflbm import Case, Simulation
from oracle_backend import OracleSimulation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("name", SMALL_CASES)
def test_fp32_two_phase_tracks_the_fp64_oracle(name, prec):
    z, ctl = load_fixture(name)
    c64 = Case(ctl, z["walls"])
    c32 = Case(ctl, z["walls"], precision=prec)
    print(f"[{prec} two-phase] case={name}", flush=True)
    g = Simulation(c32, state=c32.initial_state())
    o = OracleSimulation(c64, state=c64.initial_state())
    for s in (g, o):
        s.run(1, 40)
    mg, mo = g.macro(), o.macro()
    fl = c64.walls_global == 0
    inner = lambda a, h: a[h:-h, h:-h, h:-h][fl]
    dphi = np.abs(inner(mg["phi"], 4) - inner(mo["phi"], 4)).max()
    uscale = max(np.abs(inner(mo[k], 1)).max() for k in ("u", "v", "w"))
    du = max(np.abs(inner(mg[k], 1) - inner(mo[k], 1)).max() for k in ("u", "v", "w"))
    drho = np.abs(inner(mg["rho"], 1) - inner(mo["rho"], 1)).max()
    a, b = g.monitor(), o.monitor()
    dsat = abs(a.saturation - b.saturation)
    print(f"[{prec} two-phase] {name}: max|dphi|={dphi:.3e} max|du|={du:.3e} (u scale {uscale:.3e}) "
          f"max|drho|={drho:.3e} dsat={dsat:.3e}", flush=True)
    assert dphi < 1e-4
    assert du < 1e-3 * uscale + 1e-8
    assert drho < 1e-5
    assert dsat < 1e-6
    g.close()
