"""Closed-form properties of the step that hold independently of the reference's fixtures.

Each check is chosen so a plausible slip in the restatement (a dropped term, a wrong sign or direction
index, a transposed weight) fails it: colour-wise mass conservation (recolouring perturbation must sum to
zero), the MRT rest state (4/21, 13/126, 1/63) as a fixed point of the collision with s_e2 relaxing the
lattice-weight state, momentum j' = j + F for a uniform force, exactness of the ISO4 gradient on linear
fields, and the wetting rotation producing a unit vector at the prescribed angle to the wall normal.
"""
import math

import numpy as np

from mflbm import Case
from mflbm.case import State
from oracle_backend import OracleSimulation

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])


def box(n=12, jper=1, kper=1, **extra):
    ctl = {"lattice_dimensions": (n, n, n), "periodic_indicator": (0, jper, kper),
           "domain_wall_status_y": (0, 0) if jper else (1, 1), "domain_wall_status_z": (0, 0) if kper else (1, 1),
           "MPI_async_layers_num": (0, 4 * jper, 4 * kper), "fluid1_viscosity": 0.1, "fluid2_viscosity": 0.05,
           "surface_tension": 0.02, "theta": 90.0, "body_force_0": 0.0, "initial_fluid_distribution_option": 5,
           "initial_interface_position": 3.5}
    ctl.update(extra)
    return Case(ctl, np.zeros((n, n, n), np.int8))


def colour_mass(case, st):
    """Mass of each colour after an even step.  The AA update bounces a population off a wall over two steps
    through memory, so every wall link holds two populations in flight: the one the fluid node pushed into the
    solid node's slot at the last odd step (read back at the next odd step) and the one it left in its own slot
    toward the wall at the last even step (read back at the next even step).  Mass = sum over fluid nodes x of
    the populations the next odd update pulls, f_q(x - e_q) (periodic images sit in the halo), plus, for each
    wall link x -> x + e_r, the outgoing f_r(x)."""
    w = case.walls_global == 0
    nz, ny, nx = w.shape
    wh = case.walls_h2[1:-1, 1:-1, 1:-1]   # halo-1 view of the wall array
    out = []
    for arr in (st.f, st.g):
        tot = np.zeros_like(w, dtype=np.float64)
        for q in range(19):
            tot += arr[q, 1 - EZ[q]:1 - EZ[q] + nz, 1 - EY[q]:1 - EY[q] + ny, 1 - EX[q]:1 - EX[q] + nx]
            if q:
                nb_solid = wh[1 + EZ[q]:1 + EZ[q] + nz, 1 + EY[q]:1 + EY[q] + ny, 1 + EX[q]:1 + EX[q] + nx] == 1
                tot += np.where(nb_solid, arr[q, 1:-1, 1:-1, 1:-1], 0.0)
        out.append(tot[w].sum())
    return out


def test_colour_mass_is_conserved_in_a_periodic_box_with_a_drop():
    case = box()
    sim = OracleSimulation(case)
    sim.run(1, 2)
    m0 = colour_mass(case, sim.state())
    sim.run(3, 40)   # an even number of steps: the state is again one after an even step
    m1 = colour_mass(case, sim.state())
    assert abs(m1[0] - m0[0]) <= 1e-12 * m0[0]
    assert abs(m1[1] - m0[1]) <= 1e-12 * m0[1]


def test_mrt_rest_state_is_a_fixed_point_and_lattice_weights_relax_to_it():
    # single fluid at rest, no interface: the active MRT set relaxes the energy-square moment to 0, so the
    # rest state per direction class is rho*(4/21, 13/126, 1/63)
    case = box(initial_fluid_distribution_option=1, initial_interface_position=1000.0)
    st = case.initial_state()
    rest = np.array([4 / 21] + [13 / 126] * 6 + [1 / 63] * 12)
    sim = OracleSimulation(case, state=st)
    sim.run(1, 2)
    a = sim.state().f[:, 5, 5, 5]
    # lattice weights (1/3, 1/18, 1/36) -> distance to the rest state shrinks by (1 - s_e2) = -0.4 per step
    w0 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    np.testing.assert_allclose(a - rest, (w0 - rest) * 0.16, rtol=0, atol=1e-15)
    for q in range(19):
        st.f[q][...] = rest[q]
        st.g[q][...] = 0.0
    sim = OracleSimulation(case, state=st)
    sim.run(1, 2)
    np.testing.assert_allclose(sim.state().f[:, 5, 5, 5], rest, rtol=0, atol=1e-15)


def test_uniform_body_force_adds_momentum_f_per_step():
    force = 1e-5
    case = box(initial_fluid_distribution_option=1, initial_interface_position=1000.0, body_force_0=force)
    sim = OracleSimulation(case)
    sim.run(1, 2)
    st = sim.state()
    jz = sum(EZ[q] * st.f[q, 6, 6, 6] for q in range(19)) + sum(EZ[q] * st.g[q, 6, 6, 6] for q in range(19))
    # after two steps the post-collision momentum is 2F (one F per collision, uniform state, walls far)
    assert abs(jz - 2 * force) <= 1e-16   # rounding of order-0.1 populations


def test_iso4_gradient_is_exact_on_a_linear_field():
    case = box(n=10)
    st = case.initial_state()
    k, j, i = np.meshgrid(np.arange(-3, 15), np.arange(-3, 15), np.arange(-3, 15), indexing="ij")
    a, b, c = 0.013, -0.007, 0.011
    st.phi[...] = a * i + b * j + c * k
    sim = OracleSimulation(case, state=st, gradient=False)
    sim.color_gradient()
    gr = sim.gradient()
    norm = math.sqrt(a * a + b * b + c * c)
    sl = (slice(3, -3), slice(3, -3), slice(4, -4))   # x away from the walls at i = 1, n (extrapolated phi)
    np.testing.assert_allclose(gr["c_norm"][sl], norm, rtol=1e-12)
    np.testing.assert_allclose(gr["cn_x"][sl], a / norm, rtol=1e-12)
    np.testing.assert_allclose(gr["cn_y"][sl], b / norm, rtol=1e-12)
    np.testing.assert_allclose(gr["cn_z"][sl], c / norm, rtol=1e-12)


def test_wetting_turns_the_normal_to_the_contact_angle():
    # a flat wall at x = 1: the fluid-boundary normals point along +x; the altered interface normal must be a
    # unit vector at angle theta (radians, through fluid 2) to the wall normal
    theta_in = 60.0
    case = box(n=12, theta=theta_in)
    st = case.initial_state()
    k, j, i = np.meshgrid(np.arange(-3, 17), np.arange(-3, 17), np.arange(-3, 17), indexing="ij")
    st.phi[...] = np.tanh((j - 6.3) * 0.4 + (i - 5.0) * 0.1)
    sim = OracleSimulation(case, state=st, gradient=False)
    sim.color_gradient()
    gr = sim.gradient()
    th = case.rp.theta
    fb = case.lists.f_ijk
    nw = case.lists.f_nw
    checked = 0
    for t in range(len(fb)):
        ii, jj, kk = fb[t]
        if not (2 <= ii <= 3 and 2 <= jj <= 10 and 2 <= kk <= 10):
            continue
        v = np.array([gr["cn_x"][kk + 1, jj + 1, ii + 1], gr["cn_y"][kk + 1, jj + 1, ii + 1], gr["cn_z"][kk + 1, jj + 1, ii + 1]])
        if gr["c_norm"][kk + 1, jj + 1, ii + 1] <= 1e-6:
            continue
        assert abs(np.linalg.norm(v) - 1.0) < 1e-12
        assert abs(math.acos(np.clip(v @ nw[t], -1, 1)) - th) < 1e-9
        checked += 1
    assert checked > 10


def test_state_roundtrip_through_upload_and_download_is_exact():
    case = box(n=10)
    st = case.initial_state()
    rng = np.random.default_rng(5)
    st.f[...] = rng.random(st.f.shape)
    st.g[...] = rng.random(st.g.shape)
    sim = OracleSimulation(case, state=st, gradient=False)
    back = sim.state()
    assert np.array_equal(back.f, st.f) and np.array_equal(back.g, st.g)


def test_initial_state_is_the_lattice_weight_equilibrium():
    case = box(n=10)
    st = case.initial_state()
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    phi = st.phi[3:-3, 3:-3, 3:-3]
    rho1 = (1.0 + phi) * 0.5
    for q in range(19):
        np.testing.assert_array_equal(st.f[q], rho1 * w[q])


def test_state_class_shapes():
    s = State.empty(4, 5, 6)
    assert s.f.shape == (19, 8, 7, 6) and s.phi.shape == (14, 13, 12)
