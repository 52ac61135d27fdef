"""Boundary-node lists built on the GPU (mflbm_geometry_new_device) equal the host routine's, element for element,
on the fixture geometries (periodic, open, y/z-periodic) and on a 64^3 sphere pack."""
import numpy as np
import pytest

from helpers import SMALL_CASES, load_fixture
from mflbm import Case, geometry

pytestmark = pytest.mark.gpu


def _same_lists(a: Case, b: Case):
    for name in ("s_ijk", "s_nfl", "s_nbr", "s_w", "f_ijk", "f_nw", "f_theta"):
        x, y = getattr(a.lists, name), getattr(b.lists, name)
        assert x.shape == y.shape, name
        assert np.array_equal(x, y), name


@pytest.mark.parametrize("name", SMALL_CASES)
def test_device_geometry_matches_host_on_fixtures(name):
    z, ctl = load_fixture(name)
    print(f"[geometry] case={name}", flush=True)
    _same_lists(Case(ctl, z["walls"]), Case(ctl, z["walls"], geometry="device"))


def test_device_geometry_matches_host_on_a_sphere_pack():
    walls = geometry.sphere_pack(64, 64, 64, 0.3, 3.0, 6.0, 5, buffer=4)
    ctl = {"lattice_dimensions": (64, 64, 64), "periodic_indicator": (0, 0, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 0, 4), "theta": 120.0}
    print("[geometry] sphere pack 64^3", flush=True)
    _same_lists(Case(ctl, walls), Case(ctl, walls, geometry="device"))
