"""Batched single-phase runs (mflbm_spb_*) against the single-sample path.

The batch runs the same per-node update as mflbm_sp_* (which matches the reference's checkpoints bit for bit
in FP64), on samples of uneven size packed into one launch.  Checks: every sample's populations equal the
single-sample run bit for bit, in FP64 and in FP32; steady runs stop each sample at the step the
single-sample run stops at, with the same flow rate to rounding of the reduction order; a stopped sample keeps
its state; FP32 permeabilities are within 0.5 % of FP64.
"""
import numpy as np
import pytest

from mflbm import geometry
from mflbm.batch import REVBatch, REVSample, area_xy, global_walls
from mflbm.singlephase import SPCase, SPSimulation, permeability

pytestmark = pytest.mark.gpu

NU, FORCE = 0.1, 1e-5


def _samples():
    out = []
    for k, (n, phi, jper) in enumerate([(20, 0.35, 0), (28, 0.3, 1), (24, 0.25, 0), (33, 0.3, 0)]):
        w = geometry.sphere_pack(n, n + 2, n + 4, phi, 3.0, 5.0, 40 + k, buffer=3)
        out.append(REVSample(w, jper))
    return out


def _case(s: REVSample, precision: str, tol: float = 1e-6) -> SPCase:
    nx, ny, nz = s.shape
    ctl = {"lattice_dimensions": (nx, ny, nz), "periodic_indicator": (0, s.jper, 1),
           "domain_wall_status_y": (0, 0) if s.jper else (1, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 4 if s.jper else 0, 4), "fluid_viscosity": NU, "body_force_0": FORCE,
           "MRT_collision_parameter_preset": 1, "steady_state_option": 1, "convergence_criteria": tol}
    return SPCase(ctl, s.walls, precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_batch_matches_single_sample_runs_bit_for_bit(precision):
    samples = _samples()
    print(f"[sp batch parity] precision={precision} samples={[s.shape for s in samples]}", flush=True)
    b = REVBatch(samples, NU, FORCE, precision, block=128, tile_rows=16)
    b.run(1, 37)
    for i, s in enumerate(samples):
        sim = SPSimulation(_case(s, precision))
        sim.run(1, 37)
        fs = sim.state().f
        fb = b.state(i)
        fl = global_walls(s) == 0
        assert b.sample_fluid(i) == int(fl.sum())
        for q in range(19):
            assert np.array_equal(fb[q, 1:-1, 1:-1, 1:-1][fl], fs[q, 1:-1, 1:-1, 1:-1][fl]), (i, q)
        sim.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_rest_state_start_matches_the_collision_equilibrium_state(precision):
    """rest_state = 1: every sample starts at the collision's rest equilibrium, as a single-sample run started
    from SPCase.collision_equilibrium_state()."""
    samples = _samples()[:2]
    print(f"[sp batch rest state] precision={precision}", flush=True)
    b = REVBatch(samples, NU, FORCE, precision, rest_state=1)
    b.run(1, 23)
    for i, s in enumerate(samples):
        case = _case(s, precision)
        sim = SPSimulation(case, case.collision_equilibrium_state())
        sim.run(1, 23)
        fs, fb = sim.state().f, b.state(i)
        fl = global_walls(s) == 0
        for q in range(19):
            assert np.array_equal(fb[q, 1:-1, 1:-1, 1:-1][fl], fs[q, 1:-1, 1:-1, 1:-1][fl]), (i, q)
        sim.close()
    b.close()


def test_batch_steady_runs_stop_where_the_single_sample_runs_stop():
    samples = _samples()
    tol, nm = 1e-6, 100
    b = REVBatch(samples, NU, FORCE, "fp64", ntime_monitor=nm, tol=tol)
    r = b.run_steady(40000)
    print(f"[sp batch steady] conv={r.conv_step.tolist()} periods={r.periods} k={r.permeability.tolist()}",
          flush=True)
    assert np.all(r.conv_step > 0)
    for i, s in enumerate(samples):
        case = _case(s, "fp64", tol)
        sim = SPSimulation(case)
        n, status, hist = sim.run_steady(1, 40000, nm)
        assert status == "converged"
        assert r.conv_step[i] == n, (i, r.conv_step[i], n)
        q = hist[-1][1]
        assert abs(r.flowrate[i] - q) <= 1e-12 * abs(q)
        assert abs(r.permeability[i] - permeability(case, q)) <= 1e-12 * abs(r.permeability[i])
        assert case.rp.A_xy == area_xy(*s.shape[:2])
        sim.close()
    # a stopped sample keeps its state while the batch runs on
    f0 = b.state(0)
    b.run(1, 10)
    assert np.array_equal(b.state(0), f0)
    b.close()


def test_batch_fp32_storage_permeability_within_half_a_percent_of_fp64():
    samples = _samples()
    res = {}
    for prec in ("fp64", "mixed"):
        b = REVBatch(samples, NU, FORCE, prec, ntime_monitor=200, tol=1e-7)
        res[prec] = b.run_steady(60000)
        assert np.all(res[prec].conv_step > 0), (prec, res[prec].conv_step)
        b.close()
    rel = np.abs(res["mixed"].permeability / res["fp64"].permeability - 1.0)
    print(f"[sp batch mixed] rel err per sample {rel.tolist()} steps fp64 {res['fp64'].conv_step.tolist()} "
          f"mixed {res['mixed'].conv_step.tolist()}", flush=True)
    assert np.all(rel < 5e-3)
