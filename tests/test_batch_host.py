"""Host side of the batched single-phase path: the synthetic REV set, the reference's cross-section area,
the autotuner's sample subset, and the library's refusal without a GPU."""
import ctypes as C

import numpy as np
import pytest

from mflbm import io, native
from mflbm.batch import REVBatch, _subset, area_xy, rev_samples


def test_rev_samples_are_seeded_and_within_the_requested_ranges():
    a = rev_samples(6, seed=11, sides=(16, 24), porosity=(0.15, 0.35), radius=(2.0, 4.0), buffer=2)
    b = rev_samples(6, seed=11, sides=(16, 24), porosity=(0.15, 0.35), radius=(2.0, 4.0), buffer=2)
    for x, y in zip(a, b):
        assert x.shape == y.shape and np.array_equal(x.walls, y.walls)
    for s in a:
        n = s.shape[0]
        assert s.shape == (n, n, n) and n in (16, 24)
        core = s.walls[2:-2, 1:-1, 1:-1]                 # the sample region between the fluid buffers
        assert not s.walls[:2].any() and not s.walls[-2:].any()
        assert 0.15 - 0.06 <= 1.0 - core.mean() <= 0.35 + 0.01
    c = rev_samples(6, seed=12, sides=(16, 24), radius=(2.0, 4.0), buffer=2)
    assert any(x.shape != y.shape or not np.array_equal(x.walls, y.walls) for x, y in zip(a, c))


@pytest.mark.parametrize("nx,ny", [(48, 48), (64, 96), (17, 33)])
def test_area_is_the_reference_cross_section(nx, ny):
    rp = io.derive_sp_params({"lattice_dimensions": (nx, ny, 8), "periodic_indicator": (0, 0, 1),
                              "domain_wall_status_z": (0, 0), "fluid_viscosity": 0.1})
    assert area_xy(nx, ny) == rp.A_xy


def test_autotune_subset_keeps_a_few_samples_per_shape():
    s = rev_samples(12, seed=3, sides=(16, 24), radius=(2.0, 4.0), buffer=2)
    sub = _subset(s, 2)
    shapes = [x.shape for x in sub]
    assert all(shapes.count(sh) <= 2 for sh in shapes)
    assert set(shapes) == {x.shape for x in s}


def test_batch_abi_is_exported_and_refuses_to_run_without_a_device():
    lib = native.load()
    for n in native._SPB_ENTRY_POINTS:
        assert hasattr(lib, "mflbm_spb_" + n)
    fn = native.bind(lib, "mflbm_", batch=True)
    p = native.SPBParams()
    p.ntime_monitor, p.la_nu, p.precision = 3, 0.1, 0
    ctx = C.c_void_p()
    assert fn["create"](C.byref(p), C.byref(ctx)) != 0           # odd monitor period
    assert b"even" in lib.mflbm_create_error()
    p.ntime_monitor = 4
    rc = fn["create"](C.byref(p), C.byref(ctx))
    if rc == 0:      # a device is visible: nothing more to check here
        fn["destroy"](ctx)
        return
    assert b"no HIP device" in lib.mflbm_create_error()
    with pytest.raises(RuntimeError):
        REVBatch(rev_samples(1, seed=1, sides=(16,), radius=(2.0, 3.0), buffer=2))
