"""Single-phase oracle (oracle/oracle_sp.cpp) pinned to the reference's single-phase solver.

* bit-exact against checkpoints of the unmodified reference (tests/golden/sp_*.npz, written by
  make_fixtures_sp.py): initial state, steps 1, 2, 3 and 40, on a periodic duct, a velocity-inlet /
  convective-outlet pack, a Zou-He pressure tube (SRT rates) and a y/z-periodic box;
* the monitor files of the reference (flow rate, umax, pressures, printed with 7-8 digits);
* closed forms: plane Poiseuille flow between bounce-back walls is exact for the bounce-back-optimised MRT
  set (Lambda = 3/16), and the SRT set shows the known slip u_s = F (16 Lambda - 3) / (24 nu); colour-free
  mass conservation in a periodic box.
"""
import numpy as np
import pytest

from helpers import SP_CASES, assert_sp_matches, fixture_steps, load_sp
from mflbm import io
from mflbm.singlephase import SPCase, duct_flow_rate, permeability
from oracle_backend import SPOracleSimulation

@pytest.mark.parametrize("name", SP_CASES)
def test_sp_oracle_trajectory_matches_reference(name):
    z, case = load_sp(name)
    st = case.initial_state()
    assert_sp_matches(z, case, st, 0)
    sim = SPOracleSimulation(case, state=st)
    done = 0
    for n in fixture_steps(z)[1:]:
        sim.run(done + 1, n - done)
        done = n
        assert_sp_matches(z, case, sim.state(), n)


def _parse(text):
    return [line.split() for line in text.strip().splitlines()]


def fortran_e(x: float, d: int) -> str:
    """Fortran Ew.d edit of x without the field padding: 0.<d digits>E<sign><2 digits>, correctly rounded."""
    if x == 0.0:
        return "0." + "0" * d + "E+00"
    m, e = f"{abs(x):.{d - 1}e}".split("e")
    digits = m.replace(".", "")
    ex = int(e) + 1
    return ("-" if x < 0 else "") + "0." + digits + f"E{'+' if ex >= 0 else '-'}{abs(ex):02d}"


@pytest.mark.parametrize("name", SP_CASES)
def test_sp_oracle_monitor_matches_reference_files(name):
    z, case = load_sp(name)
    ref = _parse(str(z["monitor_flowrate_time"]))
    pre = _parse(str(z["monitor_pre"])) if "monitor_pre" in z.files else None
    sim = SPOracleSimulation(case)
    done = 0
    for row_i, row in enumerate(ref):
        n = int(row[0])
        sim.run(done + 1, n - done)
        done = n
        m = sim.monitor()
        # flowrate_time.dat: (I10, 2(1x, E14.6)); pre.dat: (I10, 3(1x, e14.7)) (Monitor.F90:119-130)
        assert [fortran_e(m.flowrate, 6), fortran_e(m.umax, 6)] == row[1:3]
        if pre is not None:
            p = pre[row_i]
            assert int(p[0]) == n
            assert [fortran_e(v, 7) for v in m.pressure] == p[1:4]


def _plane_channel(nx, preset, nu=0.1, force=1e-5):
    ny = nz = 4
    ctl = {"lattice_dimensions": (nx, ny, nz), "periodic_indicator": (0, 1, 1), "domain_wall_status_y": (0, 0),
           "domain_wall_status_z": (0, 0), "MPI_async_layers_num": (0, 4, 4), "fluid_viscosity": nu,
           "body_force_0": force, "MRT_collision_parameter_preset": preset, "steady_state_option": 1,
           "convergence_criteria": 1e-13}
    return SPCase(ctl, np.zeros((nz, ny, nx), np.int8))


@pytest.mark.parametrize("preset", [1, 0])
def test_sp_plane_poiseuille_closed_form(preset):
    """Walls at i = 1 and i = nx sit half-way to the first fluid nodes, so the channel width is H = nx - 2 and
    node i is at x = i - 1.5.  Steady MRT with bounce-back walls gives u(x) = F/(2 nu) x (H - x) + u_s, a slip
    u_s = F (16 Lambda - 3) / (24 nu) with Lambda = (1/s_nu - 1/2)(1/s_q - 1/2) (Ginzburg & d'Humieres 2003);
    the preset-1 rates make Lambda = 3/16 and the profile exact."""
    nx, nu, force = 10, 0.1, 1e-5
    case = _plane_channel(nx, preset, nu, force)
    sim = SPOracleSimulation(case)
    n, status, hist = sim.run_steady(1, 100000, 200)
    assert status == "converged"
    omega = 1.0 / (3.0 * nu + 0.5)
    s_q = 8.0 * (2.0 - omega) / (8.0 - omega) if preset == 1 else omega
    lam = (1.0 / omega - 0.5) * (1.0 / s_q - 0.5)
    H = nx - 2
    x = np.arange(1, nx + 1) - 1.5
    u_s = force * (16.0 * lam - 3.0) / (24.0 * nu)
    wa = force / (2.0 * nu) * x * (H - x) + u_s
    wa[0] = wa[-1] = 0.0
    w = sim.macro()["w"][2, 2, 1:-1]
    np.testing.assert_allclose(w, wa, rtol=0, atol=1e-11 * wa.max())
    if preset == 1:
        assert abs(lam - 3.0 / 16.0) < 1e-15
    # flow rate: sum over the layer, mean over layers
    q = hist[-1][1]
    assert q == pytest.approx(case.rp.ny * wa.sum(), rel=1e-10)


def test_sp_duct_flow_rate_approaches_the_series_at_second_order():
    """Square-ish duct (walls in x and y), body force: the error against the analytic series shrinks about
    4x when the resolution doubles (second-order accuracy of the scheme)."""
    errs = []
    for nx, ny in ((8, 9), (14, 16)):
        ctl = {"lattice_dimensions": (nx, ny, 4), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
               "MPI_async_layers_num": (0, 0, 4), "fluid_viscosity": 0.1, "body_force_0": 1e-5,
               "MRT_collision_parameter_preset": 1, "steady_state_option": 1, "convergence_criteria": 1e-11}
        case = SPCase(ctl, np.zeros((4, ny, nx), np.int8))
        sim = SPOracleSimulation(case)
        _, status, hist = sim.run_steady(1, 100000, 200)
        assert status == "converged"
        qa = duct_flow_rate(1e-5, 0.1, nx - 2, ny - 2)
        errs.append(abs(hist[-1][1] / qa - 1.0))
        k = permeability(case, hist[-1][1])
        assert k == pytest.approx(0.1 * hist[-1][1] / (case.rp.A_xy * 1e-5), rel=1e-15)
    assert errs[0] < 0.05
    assert 2.5 < errs[0] / errs[1] < 6.0


EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])


def sp_mass(case, f):
    """Mass after an even step: the populations the next odd update pulls into fluid nodes, f_q(x - e_q)
    (periodic images sit in the halo), plus one population per wall link x -> x + e_r, the outgoing f_r(x)
    (the AA update keeps a second population of each wall link in flight in the solid node's slot)."""
    w = case.walls_global == 0
    nz, ny, nx = w.shape
    wh = case.walls_h2[1:-1, 1:-1, 1:-1]
    tot = np.zeros(w.shape)
    for q in range(19):
        tot += f[q, 1 - EZ[q]:1 - EZ[q] + nz, 1 - EY[q]:1 - EY[q] + ny, 1 - EX[q]:1 - EX[q] + nx]
        if q:
            nb = wh[1 + EZ[q]:1 + EZ[q] + nz, 1 + EY[q]:1 + EY[q] + ny, 1 + EX[q]:1 + EX[q] + nx] == 1
            tot += np.where(nb, f[q, 1:-1, 1:-1, 1:-1], 0.0)
    return tot[w].sum()


def test_sp_mass_is_conserved_in_a_periodic_box():
    z, case = load_sp("sp_periodic_yz")
    sim = SPOracleSimulation(case)
    sim.run(1, 2)
    m0 = sp_mass(case, sim.state().f)
    sim.run(3, 40)
    m1 = sp_mass(case, sim.state().f)
    assert abs(m1 - m0) <= 1e-12 * m0


def test_sp_control_and_checkpoint_roundtrip(tmp_path):
    ctl = {"lattice_dimensions": (8, 9, 10), "fluid_viscosity": 0.07, "Reynolds_number": 3.5,
           "MRT_collision_parameter_preset": 0, "excluded_layers": (2, 3)}
    p = tmp_path / "simulation_control.txt"
    io.write_control_sp(str(p), ctl)
    back = io.read_control_sp(str(p))
    for k, v in ctl.items():
        assert back[k] == v
    rng = np.random.default_rng(3)
    ck = io.SPCheckpoint(41, 1e-5, 1.01, rng.random((19, 12, 11, 10)), rng.random((19, 11, 10)))
    io.write_checkpoint_sp(str(tmp_path / "id0000"), ck)
    r = io.read_checkpoint_sp(str(tmp_path / "id0000"), 8, 9, 10, True)
    assert r.ntime_next == 41 and r.force_z == 1e-5 and r.rho_in == 1.01
    assert np.array_equal(r.f, ck.f) and np.array_equal(r.f_convec, ck.f_convec)


def test_sp_derived_constants_follow_the_reference():
    rp = io.derive_sp_params({"lattice_dimensions": (20, 30, 40), "periodic_indicator": (0, 0, 0),
                              "domain_wall_status_z": (0, 0), "inlet_BC": 1, "outlet_BC": 1,
                              "Reynolds_number": 10.0, "char_length": 20.0, "fluid_viscosity": 0.05})
    assert rp.uin_avg == 10.0 * 0.05 / 20.0
    assert rp.A_xy == 18.0 * 28.0 and rp.flowrate0 == rp.uin_avg * rp.A_xy
    assert rp.force_z == 1e-6   # the body force stays on with a velocity inlet in the single-phase program
    rp = io.derive_sp_params({"periodic_indicator": (0, 0, 0), "domain_wall_status_z": (0, 0), "inlet_BC": 2,
                              "outlet_BC": 2, "rho_drop": 0.02})
    assert rp.rho_in == 1.02 and rp.rho_out == 1.0
    with pytest.raises(ValueError):
        io.derive_sp_params({"periodic_indicator": (0, 0, 0), "domain_wall_status_z": (0, 0), "inlet_BC": 2,
                             "outlet_BC": 1})
