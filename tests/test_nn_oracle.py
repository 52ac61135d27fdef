"""The warm-start network's CPU oracle (oracle/nn_oracle.py) pinned independently of the GPU kernels:
the convolution against scipy's n-dimensional correlation, its gradients and the whole network's gradients
against central finite differences (FP32 mode), the BF16 rounding against numpy's own float32 bit layout."""
import os
import sys

import numpy as np
import pytest
import scipy.ndimage

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import nn_oracle as O  # noqa: E402


def test_bf16_rounding_is_round_to_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265], np.float32)
    r = O.to_bf16(a)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even
    assert r[2] == 1.0
    assert r[3] == 1.0 + 2 ** -7
    assert r[4] == -2.5
    assert abs(r[5] - 3.140625) < 1e-7
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_conv3d_is_the_correlation_with_zero_padding():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 5, 6, 7, 1)).astype(np.float32)
    k = rng.standard_normal((3, 3, 3)).astype(np.float32)
    w = k.reshape(1, 27, 1)                       # off = 9 (dz+1) + 3 (dy+1) + (dx+1)
    out = O.conv3d(x, w)[0, ..., 0]
    ref = scipy.ndimage.correlate(x[0, ..., 0].astype(np.float64), k.astype(np.float64), mode="constant", cval=0.0)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-5)


def _fd(f, a, idx, h=1e-2):
    a[idx] += h
    fp = f()
    a[idx] -= 2 * h
    fm = f()
    a[idx] += h
    return (fp - fm) / (2 * h)


def test_conv3d_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 4, 4, 3)).astype(np.float64)
    w = rng.standard_normal((2, 27, 3)).astype(np.float64)
    gy = rng.standard_normal((1, 4, 4, 4, 2)).astype(np.float64)
    f = lambda: float((O.conv3d(x, w).astype(np.float64) * gy).sum())
    gx, gw, gb = O.conv3d_grads(x, w, gy)
    for idx in [(0, 0, 0, 0, 0), (0, 1, 2, 3, 2), (0, 3, 3, 0, 1)]:
        assert abs(_fd(f, x, idx) - gx[idx]) < 1e-3 * (1 + abs(gx[idx]))
    for idx in [(0, 0, 0), (1, 13, 2), (1, 26, 1), (0, 5, 0)]:
        assert abs(_fd(f, w, idx) - gw[idx]) < 1e-3 * (1 + abs(gw[idx]))


def _tiny_net(bf16):
    rng = np.random.default_rng(3)
    levels, cin, cout = [8, 16], 8, 8
    params = {}
    for name, ci, co, _ in O.layer_specs(levels, cin, cout):
        params[name] = (rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci)),
                        rng.standard_normal(co) * 0.1 + 0.05)
    x = np.zeros((1, 4, 4, 8, cin))
    x[..., 0] = rng.random((1, 4, 4, 8)) > 0.3
    t = rng.standard_normal((1, 4, 4, 8, cout))
    return O.UNetOracle(params, levels, cin, cout, bf16=bf16), params, x, t


def test_network_gradients_match_finite_differences():
    net, params, x, t = _tiny_net(False)
    loss, _, g = net.loss_and_grads(x, t, 3)
    f = lambda: net.loss_and_grads(x, t, 3)[0]
    checked = 0
    for name in ("head", "dec0b", "dec0a", "enc1b", "enc1a", "enc0b", "enc0a"):
        w, b = params[name]
        gw, gb = g[name]
        for idx in [(0, 13, 0), (w.shape[0] - 1, 4, w.shape[2] - 1)]:
            fd = _fd(f, w, idx, 1e-5)   # FP64; small steps keep clear of the ReLU kinks
            assert abs(fd - gw[idx]) < 1e-6 * (abs(fd) + 1e-3), (name, idx, fd, gw[idx])
            checked += 1
        fd = _fd(f, b, (0,), 1e-5)
        assert abs(fd - gb[0]) < 1e-6 * (abs(fd) + 1e-3), (name, fd, gb[0])
    assert checked == 14


def test_bf16_network_stays_close_to_the_fp32_network():
    n16, _, x, t = _tiny_net(True)
    n32, _, _, _ = _tiny_net(False)
    y16 = n16.forward(x)[0]
    y32 = n32.forward(x)[0]
    assert np.abs(y16 - y32).max() < 5e-2 * (np.abs(y32).max() + 1e-6)
