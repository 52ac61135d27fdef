"""CPU oracle (numpy) of the warm-start network: TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the product path runs the HIP kernels of mf-lbm_amd/csrc/nn_kernels.hip.
It restates, in FP32 with BF16 rounding at the points where the device stores BF16, the operations of the
encoder-decoder (mf-lbm_amd/mflbm/warmstart.py) and their gradients:

* ``conv3d``: 3x3x3 convolution, stride 1, zero padding 1, channels-last x[n, z, y, x, c], weights
  w[co, off, ci] with off = 9 (dz+1) + 3 (dy+1) + (dx+1): out[p, co] = sum_{off, ci} x[p + d(off), ci] w[co, off, ci].
  Pinned by tests against scipy.ndimage.correlate (single channel) and by finite differences (gradients).
* ``conv3d_grads``: input gradient (the same convolution with the kernel flipped and channels swapped) and
  weight gradient.
* ``pool2`` / ``up2``: 2x average pooling and nearest upsampling with their adjoints.
* ``UNetOracle``: the network (encoder of ``levels`` channel counts, each level two ReLU convolutions,
  decoder with nearest upsampling, a convolution, a skip addition and a second convolution, a linear head)
  forward and backward for the MSE loss over the first ``used`` output channels.
There is no reference counterpart (the reference has no learned component).
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np


def to_bf16(a: np.ndarray) -> np.ndarray:
    """Round FP32 to the nearest BF16 (ties to even), returned as FP32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


OFFS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shift(xp: np.ndarray, dz: int, dy: int, dx: int, shape) -> np.ndarray:
    _, d, h, w, _ = shape
    return xp[:, 1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w, :]


def conv3d(x: np.ndarray, w: np.ndarray, b=None, relu: bool = False) -> np.ndarray:
    """x (n, d, h, w, cin), w (cout, 27, cin) -> (n, d, h, w, cout); FP32 accumulation (FP64 for FP64 inputs)."""
    dt = np.result_type(x, w, np.float32)
    xp = np.pad(x.astype(dt), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros(x.shape[:4] + (w.shape[0],), dt)
    for o, (dz, dy, dx) in enumerate(OFFS):
        out += _shift(xp, dz, dy, dx, x.shape) @ w[:, o, :].T.astype(dt)
    if b is not None:
        out += b.astype(dt)
    if relu:
        out = np.maximum(out, 0.0)
    return out


def conv3d_grads(x: np.ndarray, w: np.ndarray, gy: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Gradients of conv3d (before any activation) for output gradient gy: (gx, gw, gb)."""
    dt = np.result_type(x, w, gy, np.float32)
    wflip = np.ascontiguousarray(w[:, ::-1, :].transpose(2, 1, 0))   # (cin, 27, cout), off -> 26 - off
    gx = conv3d(gy, wflip)
    xp = np.pad(x.astype(dt), ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    gw = np.zeros(w.shape, dt)
    g2 = gy.reshape(-1, gy.shape[-1]).astype(dt)
    for o, (dz, dy, dx) in enumerate(OFFS):
        gw[:, o, :] = g2.T @ _shift(xp, dz, dy, dx, x.shape).reshape(-1, x.shape[-1])
    gb = g2.sum(axis=0)
    return gx, gw, gb


def pool2(x: np.ndarray) -> np.ndarray:
    n, d, h, w, c = x.shape
    return x.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4, 6))


def pool2_adj(g: np.ndarray) -> np.ndarray:
    return up2(g) * 0.125


def up2(x: np.ndarray) -> np.ndarray:
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def up2_adj(g: np.ndarray) -> np.ndarray:
    n, d, h, w, c = g.shape
    return g.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4, 6))


def layer_specs(levels: List[int], cin: int, cout: int) -> List[Tuple[str, int, int, bool]]:
    """(name, cin, cout, relu) of every convolution in execution order."""
    specs = []
    c = cin
    for l, ch in enumerate(levels):
        specs += [(f"enc{l}a", c, ch, True), (f"enc{l}b", ch, ch, True)]
        c = ch
    for l in range(len(levels) - 2, -1, -1):
        specs += [(f"dec{l}a", levels[l + 1], levels[l], True), (f"dec{l}b", levels[l], levels[l], True)]
    specs.append(("head", levels[0], cout, False))
    return specs


class UNetOracle:
    def __init__(self, params: Dict[str, Tuple[np.ndarray, np.ndarray]], levels: List[int], cin: int, cout: int,
                 bf16: bool = True):
        """bf16=False drops the BF16 rounding (FP64 throughout), for finite-difference checks of the gradients."""
        self.p = params          # name -> (w (cout, 27, cin) FP32 master, b (cout,) FP32)
        self.levels = levels
        self.specs = layer_specs(levels, cin, cout)
        self.r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))

    def _conv(self, name, x, relu):
        w, b = self.p[name]
        return self.r(conv3d(x, self.r(w), b, relu))

    def forward(self, x: np.ndarray):
        acts = {"in": x}
        h = x
        skips = []
        L = len(self.levels)
        for l in range(L):
            acts[f"enc{l}a"] = h = self._conv(f"enc{l}a", h, True)
            acts[f"enc{l}b"] = h = self._conv(f"enc{l}b", h, True)
            skips.append(h)
            if l < L - 1:
                acts[f"pool{l}"] = h = self.r(pool2(h))
        for l in range(L - 2, -1, -1):
            acts[f"up{l}"] = h = self.r(up2(h))
            acts[f"dec{l}a"] = h = self._conv(f"dec{l}a", h, True)
            acts[f"skip{l}"] = h = self.r(h + skips[l])
            acts[f"dec{l}b"] = h = self._conv(f"dec{l}b", h, True)
        acts["head"] = y = self._conv("head", h, False)
        return y, acts

    def loss_and_grads(self, x: np.ndarray, t: np.ndarray, used: int):
        """Loss = sum over positions and the first ``used`` channels of (y - t)^2, divided by the number of
        positions; returns (loss, output, {name: (gw, gb)})."""
        y, A = self.forward(x)
        npos = int(np.prod(y.shape[:4]))
        d = np.zeros_like(y)
        d[..., :used] = y[..., :used] - t[..., :used]
        loss = float((d.astype(np.float64) ** 2).sum()) / npos
        g = self.r(2.0 * (1.0 / npos) * d)
        grads = {}
        L = len(self.levels)

        def back(name, xin, yout, g, relu):
            if relu:
                g = np.where(yout > 0, g, 0.0).astype(g.dtype)
            w, _ = self.p[name]
            gx, gw, gb = conv3d_grads(xin, self.r(w), g)
            grads[name] = (gw, gb)
            return self.r(gx)

        top = A["dec0b"] if L > 1 else A["enc0b"]
        g = back("head", top, A["head"], g, False)
        gskip = {}
        for l in range(0, L - 1):        # decoder levels, the reverse of their forward order
            g = back(f"dec{l}b", A[f"skip{l}"], A[f"dec{l}b"], g, True)
            gskip[l] = g
            g = back(f"dec{l}a", A[f"up{l}"], A[f"dec{l}a"], g, True)
            g = self.r(up2_adj(g))       # into the coarser level's output
        for l in range(L - 1, -1, -1):   # encoder levels, coarsest first
            g = back(f"enc{l}b", A[f"enc{l}a"], A[f"enc{l}b"], g, True)
            xin = A[f"pool{l - 1}"] if l > 0 else A["in"]
            g = back(f"enc{l}a", xin, A[f"enc{l}a"], g, True)
            if l > 0:
                g = self.r(self.r(pool2_adj(g)) + gskip[l - 1])
        return loss, y, grads
