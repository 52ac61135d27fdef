"""The allocation-free weight gradient (mflbm_nn_conv3d_dw_ordered_noalloc) against the workspace form
(mflbm_nn_conv3d_dw_ordered), through the C ABI: identical bits on two calls in a row, the second smaller than the
first (the scratch array keeps the first call's contents, which the second must not read), both accumulating onto
gradients that already hold values; a size beyond the scratch array is refused before any launch."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import nn_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(cin, cout, shape, seed):
    import torch
    rng = np.random.default_rng(seed)
    x = O.to_bf16(rng.standard_normal(shape + (cin,)).astype(np.float32))
    g = O.to_bf16(rng.standard_normal(shape + (cout,)).astype(np.float32))
    dw0 = rng.standard_normal((cout, 27 * cin)).astype(np.float32)
    db0 = rng.standard_normal(cout).astype(np.float32)
    dev = lambda a, t: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", t)
    return dev(x, torch.bfloat16), dev(g, torch.bfloat16), dev(dw0, torch.float32), dev(db0, torch.float32)


def test_allocation_free_weight_gradient_equals_the_workspace_form():
    import torch
    from mflbm import native
    lib = native.load()
    P = lambda t: C.c_void_p(t.data_ptr())
    cap = lib.mflbm_nn_conv3d_dw_scratch_floats()
    assert cap == 4 << 20
    # 18432 positions (5 weight chunks, 3 bias chunks), then 480 positions (one chunk each)
    for cin, cout, shape, seed in ((16, 48, (2, 12, 16, 48), 5), (8, 32, (2, 6, 5, 8), 6)):
        print(f"[no-alloc dw] cin {cin} cout {cout} shape {shape}", flush=True)
        need = lib.mflbm_nn_conv3d_dw_work(*shape, cin, cout)
        assert 0 < need <= cap
        x, g, dw0, db0 = _case(cin, cout, shape, seed)
        work = torch.empty(need, device="cuda")
        dw_ref, db_ref = dw0.clone(), db0.clone()
        assert lib.mflbm_nn_conv3d_dw_ordered(P(x), P(g), P(dw_ref), P(db_ref), P(work), need, *shape, cin, cout,
                                              None) == 0
        dw_new, db_new = dw0.clone(), db0.clone()
        assert lib.mflbm_nn_conv3d_dw_ordered_noalloc(P(x), P(g), P(dw_new), P(db_new), *shape, cin, cout, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(dw_new, dw_ref) and torch.equal(db_new, db_ref)
        assert not torch.equal(dw_new, dw0)
        # without the bias gradient: the weight gradient alone, same bits
        dw_nb = dw0.clone()
        assert lib.mflbm_nn_conv3d_dw_ordered_noalloc(P(x), P(g), P(dw_nb), None, *shape, cin, cout, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(dw_nb, dw_ref)


def test_allocation_free_weight_gradient_refuses_sizes_beyond_its_scratch():
    import torch
    from mflbm import native
    lib = native.load()
    P = lambda t: C.c_void_p(t.data_ptr())
    cin = cout = 64
    shape = (1, 32, 32, 160)   # 163840 positions: 40 chunks of 64 x 27 x 64 floats, beyond 4 Mi floats
    assert lib.mflbm_nn_conv3d_dw_work(*shape, cin, cout) > lib.mflbm_nn_conv3d_dw_scratch_floats()
    x = torch.zeros(int(np.prod(shape)) * cin, dtype=torch.bfloat16, device="cuda")
    g = torch.zeros(int(np.prod(shape)) * cout, dtype=torch.bfloat16, device="cuda")
    dw = torch.zeros(cout * 27 * cin, device="cuda")
    assert lib.mflbm_nn_conv3d_dw_ordered_noalloc(P(x), P(g), P(dw), None, *shape, cin, cout, None) == 1
    torch.cuda.synchronize()
    assert float(dw.abs().max()) == 0.0
