"""Field-output files (mflbm.vtk) against the files the reference writes (tests/golden/vtk_files.npz, made by
tests/golden/make_vtk_fixtures.py from the reference binaries): the oracle's state after 10 steps of the
periodic-tube and single-phase obstacle cases, written with the same writers, must match the reference's
files byte for byte (sha256).  This also pins the macroscopic fields of compute_macro_vars bit for bit."""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, fixture_case, load_sp
from mflbm import vtk
from oracle_backend import OracleSimulation, SPOracleSimulation

Z = np.load(os.path.join(GOLDEN, "vtk_files.npz"))
FILES = json.loads(str(Z["files"]))
STEPS = int(Z["steps"])


def _sha(path):
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def _check(path, expect):
    assert os.path.getsize(path) == expect["size"]
    assert _sha(path) == expect["sha256"], os.path.basename(path)


def _inner(a, h):
    return a[h:-h, h:-h, h:-h]


@pytest.mark.parametrize("prec", [1, 0])
def test_two_phase_vtk_files_match_the_reference(tmp_path, prec):
    ref = FILES[f"periodic_tube_prec{prec}"]
    z, case = fixture_case("periodic_tube")
    p = str(tmp_path / "walls_bin.vtk")
    vtk.write_walls(p, case.walls_global)
    _check(p, ref["walls_bin.vtk"])
    st = case.initial_state()
    p = str(tmp_path / ("small_" + vtk.vtk_name(1)))
    vtk.write_phase(p, _inner(st.phi, 4), case.walls_global)
    _check(p, ref["phase_distribution/small_" + vtk.vtk_name(1)])
    sim = OracleSimulation(case, state=st)
    sim.run(1, STEPS)
    m = sim.macro()
    p = str(tmp_path / ("full_flow_field_" + vtk.vtk_name(STEPS)))
    vtk.write_full_flow(p, _inner(m["phi"], 4), _inner(m["rho"], 1), _inner(m["u"], 1), _inner(m["v"], 1),
                        _inner(m["w"], 1), double=prec == 1)
    _check(p, ref["full_flow_field/full_flow_field_" + vtk.vtk_name(STEPS)])
    p = str(tmp_path / ("small_" + vtk.vtk_name(STEPS)))
    vtk.write_phase(p, _inner(m["phi"], 4), case.walls_global)
    _check(p, ref["phase_distribution/small_" + vtk.vtk_name(STEPS)])
    back = vtk.read_vtk(str(tmp_path / ("full_flow_field_" + vtk.vtk_name(STEPS))))
    assert back["dims"] == (14, 14, 24)
    np.testing.assert_array_equal(back["velocity_Z"], _inner(m["w"], 1).astype(np.float64 if prec else np.float32))


@pytest.mark.parametrize("prec", [1, 0])
def test_single_phase_vtk_files_match_the_reference(tmp_path, prec):
    ref = FILES[f"sp_periodic_obstacle_prec{prec}"]
    z, case = load_sp("sp_periodic_obstacle")
    p = str(tmp_path / "walls_bin.vtk")
    vtk.write_walls(p, case.walls_global)
    _check(p, ref["walls_bin.vtk"])
    sim = SPOracleSimulation(case)
    sim.run(1, STEPS)
    m = sim.macro()
    p = str(tmp_path / ("full_flow_field_" + vtk.vtk_name(STEPS)))
    vtk.write_sp_full_flow(p, _inner(m["rho"], 1), _inner(m["u"], 1), _inner(m["v"], 1), _inner(m["w"], 1),
                           double=prec == 1)
    _check(p, ref["full_flow_field_" + vtk.vtk_name(STEPS)])


def test_raw_dumps_and_half_resolution_layout(tmp_path):
    a = np.arange(4 * 6 * 8, dtype=np.float64).reshape(4, 6, 8) * 0.5
    p = str(tmp_path / "phi")
    vtk.save_phi(p, a, double=True, idx=(1, 2, 3))
    raw = open(p, "rb").read()
    assert np.frombuffer(raw[:24], "<i4").tolist() == [1, 2, 3, 8, 6, 4]
    np.testing.assert_array_equal(np.frombuffer(raw[24:], "<f8").reshape(4, 6, 8), a)
    p = str(tmp_path / "half.vtk")
    vtk.write_sp_full_flow_half(p, a, a + 1, a + 2, a + 3)
    back = vtk.read_vtk(p)
    assert back["dims"] == (4, 3, 2) and back["points"] == 8 * 6 * 4 // 8
    np.testing.assert_array_equal(back["velocity_Y"], (a + 2)[::2, ::2, ::2].astype(np.float32))
