"""Boundary-info files (geometry_preprocess_cmd 1): the file the reference's preprocessing tool writes
(fixtures boundary_*.npz, made by make_boundary_fixtures.py with the unmodified tool) is read back into the
same lists the in-code geometry pass builds: identical solid lists and fluid-node coordinates, wall normals
equal up to the tool's normalisation constant (1e-16 where the solver uses eps = 1.110223025e-16), and a run
from either list set gives the same state to round-off."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, load_fixture
from mflbm import Case, io
from oracle_backend import OracleSimulation


@pytest.mark.parametrize("name", ["open_velocity", "periodic_tube"])
def test_reference_boundary_file_gives_the_in_code_lists(name, tmp_path):
    z, ctl = load_fixture(name)
    path = tmp_path / "indirect_addressing_walls_boundary.dat"
    path.write_bytes(np.load(os.path.join(GOLDEN, f"boundary_{name}.npz"))["data"].tobytes())
    a = Case(ctl, z["walls"])
    b = Case(ctl, z["walls"], boundary_file=str(path))
    for key in ("s_ijk", "s_nfl", "s_nbr", "s_w", "f_ijk", "f_theta"):
        assert np.array_equal(getattr(a.lists, key), getattr(b.lists, key)), key
    np.testing.assert_allclose(b.lists.f_nw, a.lists.f_nw, rtol=0, atol=1e-14)
    sa, sb = OracleSimulation(a), OracleSimulation(b)
    sa.run(1, 20)
    sb.run(1, 20)
    np.testing.assert_allclose(sb.state().phi, sa.state().phi, rtol=0, atol=1e-10)


def test_boundary_file_roundtrip_both_record_layouts(tmp_path):
    z, ctl = load_fixture("open_velocity")
    a = Case(ctl, z["walls"])
    L = a.lists
    nx, ny, nz = ctl["lattice_dimensions"]
    for packed in (False, True):
        p = tmp_path / f"b{int(packed)}.dat"
        io.write_boundary_info(str(p), L.s_ijk, L.s_nfl, L.s_nbr, L.s_w, L.f_ijk, L.f_nw, packed=packed)
        s_ijk, s_nfl, s_nbr, s_w, f_ijk, f_nw = io.read_boundary_info(str(p), nx, ny, nz)
        assert np.array_equal(s_ijk, L.s_ijk) and np.array_equal(s_nbr, L.s_nbr) and np.array_equal(s_w, L.s_w)
        assert np.array_equal(f_ijk, L.f_ijk) and np.array_equal(f_nw, L.f_nw)
