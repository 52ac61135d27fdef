"""Learned warm start for single-phase permeability runs (no reference counterpart: the reference starts every
run from rest, singlephase_3D/0.src/Initialization.F90:207-307).

A 3-D convolutional encoder-decoder maps a wall array to the steady velocity and density fields of the
body-force-driven flow; the fields seed the populations of a new run at the collision's own equilibrium
(``SPCase.collision_equilibrium_state``), and the unchanged solver then converges to the reference's criterion.  The network is trained on the GPU by the
project's own kernels (``mflbm_nn_*`` of include/mflbm.h: BF16 matrix-core convolutions with FP32 accumulation,
FP32 Adam master weights) on fields the batched solver produces (``mflbm/batch.py``).

Network (``layer_specs``): ``levels`` = (32, 64, 128, 256) channels; each encoder level two 3x3x3 ReLU
convolutions then 2x average pooling; each decoder level nearest 2x upsampling, a ReLU convolution, the skip
connection added, a second ReLU convolution; a linear 3x3x3 head.  Input channels (``features``): the fluid
indicator (1 fluid, 0 solid), the distance to the nearest solid node d / 8 and (d / 8)^2 (the local duct-flow
scale of the velocity), d capped at 16 and measured through the periodic faces; output channels 0..2 are (u, v, w) * nu / F (a length squared, O(1) in pores of a few
lattice units) and channel 3 is (rho - mean rho) / (P_SCALE F) (pressure differences across a pore of size d are
about 3 F d in lattice units); channels are padded to multiples of 8.  torch provides device memory and the
stream only.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import batch, native

LEVELS = (32, 64, 128, 256)
CIN = 8
COUT = 8
USED = 4
P_SCALE = 30.0


NFEAT = 3
DCAP = 16
# The network also sees the flow after T_PROBE steps of the run it is asked to start, from rest at the collision's
# equilibrium (channels NFEAT.. NFEAT + 3: u, v, w scaled as the targets, and the density deviation): the early
# response of the pressure field carries the connectivity of the pores that the wall array alone does not.
T_PROBE = 200
NIN = NFEAT + 4


def device_features(x, fluid: np.ndarray, periodic: Tuple[bool, bool, bool] = (True, False, False)) -> None:
    """Write the NFEAT input features of ``fluid`` ((nz, ny, nx) bool) into channels 0..2 of ``x``, a contiguous
    (nz, ny, nx, CIN) BF16 device tensor, with mflbm_nn_distance_features (the GPU form of ``features``)."""
    torch = _torch()
    lib = native.load()
    nz, ny, nx = fluid.shape
    fl = torch.from_numpy(np.ascontiguousarray(fluid, dtype=np.int8)).to("cuda")
    work = torch.empty(2 * fl.numel(), dtype=torch.int32, device="cuda")
    rc = lib.mflbm_nn_distance_features(_p(fl), _p(work), _p(x), nz, ny, nx, x.shape[-1], int(periodic[0]),
                                        int(periodic[1]), int(periodic[2]),
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise RuntimeError(f"mflbm_nn_distance_features failed with code {rc}")


def features(fluid: np.ndarray, periodic: Tuple[bool, bool, bool] = (True, False, False)) -> np.ndarray:
    """(nz, ny, nx) bool -> (nz, ny, nx, NFEAT) FP32: indicator, d / 8, (d / 8)^2 with d the Euclidean distance
    to the nearest solid node (0 at solids), capped at DCAP; periodic axes are wrapped by DCAP layers first.
    The host form of ``device_features`` (which the network uses), kept to check it."""
    from scipy import ndimage
    pad = [(DCAP, DCAP) if p else (0, 0) for p in periodic]
    fp = np.pad(fluid, pad, mode="wrap") if any(periodic) else fluid
    d = ndimage.distance_transform_edt(fp)
    d = d[tuple(slice(DCAP, -DCAP) if p else slice(None) for p in periodic)]
    d = np.minimum(d, DCAP) / 8.0
    return np.stack([fluid.astype(np.float32), d.astype(np.float32), (d * d).astype(np.float32)], axis=-1)


def normalized_fields(u, v, w, rho, fluid: np.ndarray, la_nu: float, force: float) -> np.ndarray:
    """Interior (nz, ny, nx, 4) FP32 of halo-1 macroscopic fields: (u, v, w) * nu / F and (rho - mean over the
    fluid) / (P_SCALE F), zero at solids."""
    t = np.stack([a[1:-1, 1:-1, 1:-1] for a in (u, v, w, rho)], axis=-1).astype(np.float64)
    t[..., :3] *= la_nu / force
    dens = t[..., 3]
    t[..., 3] = np.where(fluid, (dens - dens[fluid].mean()) / (P_SCALE * force), 0.0)
    t[~fluid] = 0.0
    return t.astype(np.float32)


def layer_specs(levels: Sequence[int], cin: int = CIN, cout: int = COUT) -> List[Tuple[str, int, int, bool]]:
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


def _kpad(c: int) -> int:
    return (27 * c + 31) // 32 * 32


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the warm-start network runs on the GPU (no HIP device visible)")
    return torch


class _Conv:
    def __init__(self, torch, name: str, cin: int, cout: int, relu: bool, gen):
        self.name, self.cin, self.cout, self.relu = name, cin, cout, relu
        k = 27 * cin
        dev = "cuda"
        self.w = (torch.randn(cout, k, generator=gen) * math.sqrt(2.0 / k)).to(dev)
        self.b = torch.zeros(cout, device=dev)
        self.gw, self.gb = torch.zeros_like(self.w), torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.wf = torch.zeros(cout, _kpad(cin), dtype=torch.bfloat16, device=dev)
        self.wb = torch.zeros(cin, _kpad(cout), dtype=torch.bfloat16, device=dev)


def _p(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


class WarmStartNet:
    """The encoder-decoder on the GPU.  ``forward`` keeps the activations the backward pass needs."""

    def __init__(self, levels: Sequence[int] = LEVELS, seed: int = 0):
        torch = _torch()
        self.torch = torch
        self.lib = native.load()
        self.levels = tuple(levels)
        gen = torch.Generator().manual_seed(seed)
        self.layers: Dict[str, _Conv] = {}
        for name, ci, co, relu in layer_specs(self.levels):
            self.layers[name] = _Conv(torch, name, ci, co, relu, gen)
        self.step_count = 0
        self._work = None   # FP32 workspace of the fixed-order gradient sums
        self.refresh()

    # -- plumbing -----------------------------------------------------------------------------------------
    def _s(self) -> C.c_void_p:
        return C.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def _ck(self, rc: int, what: str) -> None:
        if rc != 0:
            raise RuntimeError(f"{what} failed with code {rc}")

    def refresh(self) -> None:
        """BF16 forward and input-gradient weights from the FP32 master weights."""
        for L in self.layers.values():
            self._ck(self.lib.mflbm_nn_weights_to_bf16(_p(L.w), _p(L.wf), _p(L.wb), L.cin, L.cout, self._s()),
                     "mflbm_nn_weights_to_bf16")

    def _conv(self, L: _Conv, x, shape):
        n, d, h, w = shape
        y = self.torch.empty((n, d, h, w, L.cout), dtype=self.torch.bfloat16, device="cuda")
        self._ck(self.lib.mflbm_nn_conv3d_fwd(_p(x), _p(L.wf), _p(L.b), _p(y), n, d, h, w, L.cin, L.cout,
                                              1 if L.relu else 0, self._s()), "mflbm_nn_conv3d_fwd")
        return y

    def _pool(self, x, shape, c):
        n, d, h, w = shape
        y = self.torch.empty((n, d // 2, h // 2, w // 2, c), dtype=self.torch.bfloat16, device="cuda")
        self._ck(self.lib.mflbm_nn_pool2(_p(x), _p(y), n, d, h, w, c, 0, self._s()), "mflbm_nn_pool2")
        return y

    def _up(self, x, fine, c):
        n, d, h, w = fine
        y = self.torch.empty((n, d, h, w, c), dtype=self.torch.bfloat16, device="cuda")
        self._ck(self.lib.mflbm_nn_up2(_p(x), _p(y), n, d, h, w, c, 0, self._s()), "mflbm_nn_up2")
        return y

    def _add(self, a, b):
        y = self.torch.empty_like(a)
        self._ck(self.lib.mflbm_nn_add(_p(a), _p(b), _p(y), a.numel(), self._s()), "mflbm_nn_add")
        return y

    # -- forward / backward --------------------------------------------------------------------------------
    def forward(self, x):
        """x: (n, d, h, w, 8) BF16 on the device, d/h/w divisible by 2^(levels-1) and w/2^(levels-1) by 8."""
        n, d, h, w, _ = x.shape
        A = {"in": x}
        shapes = []
        hcur, shp = x, (n, d, h, w)
        Lc = len(self.levels)
        for l in range(Lc):
            shapes.append(shp)
            A[f"enc{l}a"] = hcur = self._conv(self.layers[f"enc{l}a"], hcur, shp)
            A[f"enc{l}b"] = hcur = self._conv(self.layers[f"enc{l}b"], hcur, shp)
            if l < Lc - 1:
                A[f"pool{l}"] = hcur = self._pool(hcur, shp, self.levels[l])
                shp = (shp[0], shp[1] // 2, shp[2] // 2, shp[3] // 2)
        for l in range(Lc - 2, -1, -1):
            A[f"up{l}"] = hcur = self._up(hcur, shapes[l], self.levels[l + 1])
            A[f"dec{l}a"] = hcur = self._conv(self.layers[f"dec{l}a"], hcur, shapes[l])
            A[f"skip{l}"] = hcur = self._add(hcur, A[f"enc{l}b"])
            A[f"dec{l}b"] = hcur = self._conv(self.layers[f"dec{l}b"], hcur, shapes[l])
        A["head"] = y = self._conv(self.layers["head"], hcur, shapes[0])
        self._acts, self._shapes = A, shapes
        return y

    def _back(self, name: str, xin, yout, g, shape, need_dx: bool = True):
        L = self.layers[name]
        n, d, h, w = shape
        if L.relu:
            self._ck(self.lib.mflbm_nn_relu_bwd(_p(g), _p(yout), g.numel(), self._s()), "mflbm_nn_relu_bwd")
        # weight and bias gradients summed in a fixed order: a repeated training run gives the same weights
        need = int(self.lib.mflbm_nn_conv3d_dw_work(n, d, h, w, L.cin, L.cout))
        if self._work is None or self._work.numel() < need:
            self._work = self.torch.empty(need, dtype=self.torch.float32, device="cuda")
        self._ck(self.lib.mflbm_nn_conv3d_dw_ordered(_p(xin), _p(g), _p(L.gw), _p(L.gb), _p(self._work),
                                                     self._work.numel(), n, d, h, w, L.cin, L.cout, self._s()),
                 "mflbm_nn_conv3d_dw_ordered")
        if not need_dx:
            return None
        gx = self.torch.empty((n, d, h, w, L.cin), dtype=self.torch.bfloat16, device="cuda")
        self._ck(self.lib.mflbm_nn_conv3d_fwd(_p(g), _p(L.wb), None, _p(gx), n, d, h, w, L.cout, L.cin, 0, self._s()),
                 "mflbm_nn_conv3d_fwd (input gradient)")
        return gx

    def loss_and_backward(self, x, t) -> float:
        """Forward, mean squared error over the first USED output channels (per position), backward into the
        FP32 weight gradients; returns the loss."""
        torch = self.torch
        y = self.forward(x)
        A, S = self._acts, self._shapes
        for L in self.layers.values():
            L.gw.zero_()
            L.gb.zero_()
        n, d, h, w, c = y.shape
        npos = n * d * h * w
        g = torch.empty_like(y)
        loss = torch.zeros(1, device="cuda")
        self._ck(self.lib.mflbm_nn_mse(_p(y), _p(t), _p(g), _p(loss), npos, c, USED, 1.0 / npos, self._s()),
                 "mflbm_nn_mse")
        Lc = len(self.levels)
        top = A["dec0b"] if Lc > 1 else A["enc0b"]
        g = self._back("head", top, A["head"], g, S[0])
        gskip = {}
        for l in range(0, Lc - 1):
            g = self._back(f"dec{l}b", A[f"skip{l}"], A[f"dec{l}b"], g, S[l])
            gskip[l] = g
            g = self._back(f"dec{l}a", A[f"up{l}"], A[f"dec{l}a"], g.clone(), S[l])
            coarse = (S[l][0], S[l][1] // 2, S[l][2] // 2, S[l][3] // 2, self.levels[l + 1])
            gc = torch.empty(coarse, dtype=torch.bfloat16, device="cuda")
            nn_, dd, hh, ww = S[l]
            self._ck(self.lib.mflbm_nn_up2(_p(g), _p(gc), nn_, dd, hh, ww, self.levels[l + 1], 1, self._s()),
                     "mflbm_nn_up2 (adjoint)")
            g = gc
        for l in range(Lc - 1, -1, -1):
            g = self._back(f"enc{l}b", A[f"enc{l}a"], A[f"enc{l}b"], g, S[l])
            xin = A[f"pool{l - 1}"] if l > 0 else A["in"]
            g = self._back(f"enc{l}a", xin, A[f"enc{l}a"], g, S[l], need_dx=l > 0)
            if l > 0:
                fine = S[l - 1]
                gf = torch.empty(fine + (self.levels[l - 1],), dtype=torch.bfloat16, device="cuda")
                self._ck(self.lib.mflbm_nn_pool2(_p(g), _p(gf), *fine, self.levels[l - 1], 1, self._s()),
                         "mflbm_nn_pool2 (adjoint)")
                g = self._add(gf, gskip[l - 1])
        self._acts = None
        return float(loss.item()) / npos

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.step_count += 1
        for L in self.layers.values():
            self._ck(self.lib.mflbm_nn_adam(_p(L.w), _p(L.gw), _p(L.mw), _p(L.vw), L.w.numel(), lr, b1, b2, eps,
                                            self.step_count, self._s()), "mflbm_nn_adam")
            self._ck(self.lib.mflbm_nn_adam(_p(L.b), _p(L.gb), _p(L.mb), _p(L.vb), L.b.numel(), lr, b1, b2, eps,
                                            self.step_count, self._s()), "mflbm_nn_adam")
        self.refresh()

    # -- parameters ----------------------------------------------------------------------------------------
    def get_params(self) -> Dict[str, Tuple[np.ndarray, np.ndarray]]:
        """name -> (w (cout, 27, cin), b) as FP32 numpy arrays."""
        return {n: (L.w.detach().cpu().numpy().reshape(L.cout, 27, L.cin), L.b.detach().cpu().numpy())
                for n, L in self.layers.items()}

    def set_params(self, params: Dict[str, Tuple[np.ndarray, np.ndarray]]) -> None:
        torch = self.torch
        for n, (w, b) in params.items():
            L = self.layers[n]
            L.w.copy_(torch.from_numpy(np.ascontiguousarray(w, np.float32).reshape(L.cout, -1)))
            L.b.copy_(torch.from_numpy(np.ascontiguousarray(b, np.float32)))
        self.refresh()

    def save(self, path: str) -> None:
        p = self.get_params()
        np.savez(path, levels=np.array(self.levels), **{f"{k}_w": v[0] for k, v in p.items()},
                 **{f"{k}_b": v[1] for k, v in p.items()})

    def load(self, path: str) -> None:
        z = np.load(path)
        if tuple(int(v) for v in z["levels"]) != self.levels:
            raise ValueError("saved network has other levels")
        self.set_params({n: (z[f"{n}_w"], z[f"{n}_b"]) for n in self.layers})

    # -- use ------------------------------------------------------------------------------------------------
    def predict(self, fluid: np.ndarray, periodic: Tuple[bool, bool, bool] = (True, False, False),
                probe: Optional[np.ndarray] = None) -> np.ndarray:
        """fluid indicator (nz, ny, nx) -> normalized (u, v, w, p) (4, nz, ny, nx) (FP32 numpy), zero at solids;
        ``periodic`` = (z, y, x) periodicity of the sample for the distance features; ``probe`` (nz, ny, nx, 4)
        the normalized fields after T_PROBE steps (``probe_fields``)."""
        torch = self.torch
        nz, ny, nx = fluid.shape
        x = torch.zeros((1, nz, ny, nx, CIN), dtype=torch.bfloat16, device="cuda")
        device_features(x[0], fluid > 0, periodic)
        if probe is not None:
            x[0, ..., NFEAT:NIN] = torch.from_numpy(np.ascontiguousarray(probe, np.float32)).to("cuda", torch.bfloat16)
        with torch.no_grad():
            y = self.forward(x)
        self._acts = None
        u = y[0, ..., :USED].float().cpu().numpy().transpose(3, 0, 1, 2)
        return u * (fluid[None] > 0)


# ------------------------------------------------------------------------------------------------------------
# training data from the batched solver, training loop, warm-started runs
# ------------------------------------------------------------------------------------------------------------

@dataclasses.dataclass
class Dataset:
    x: object          # (n, s, s, s, 8) BF16 on the device
    t: object          # (n, s, s, s, 8) BF16 on the device
    conv_steps: np.ndarray
    seconds: float
    target_rms: Optional[np.ndarray] = None   # per output channel, over fluid positions


def make_dataset(n: int, side: int = 64, seed: int = 11, la_nu: float = 0.1, force: float = 1e-6,
                 tol: float = 1e-6, max_steps: int = 60000, per_batch: int = 256,
                 porosity: Tuple[float, float] = (0.15, 0.35), log=None) -> Dataset:
    """n steady fields of seeded sphere packs (side^3, z-periodic, x/y walls), solved by REVBatch in the mixed
    precision mode to the flow-rate criterion ``tol``; targets (u, v, w) * nu / F and (rho - mean) / (P_SCALE F)."""
    torch = _torch()
    x = torch.zeros((n, side, side, side, CIN), dtype=torch.bfloat16, device="cuda")
    t = torch.zeros((n, side, side, side, COUT), dtype=torch.bfloat16, device="cuda")
    steps = np.zeros(n, np.int64)
    sq = np.zeros(USED)
    npos = 0
    t0 = time.perf_counter()
    done = 0
    k = 0
    while done < n:
        m = min(per_batch, n - done)
        samples = batch.rev_samples(m, seed=seed + 7919 * k, sides=(side,), porosity=porosity)
        # started at the collision's rest state: no start-up density waves, so every pore, connected or not,
        # keeps density 1 on average and the density target is the flow's own pressure field
        b = batch.REVBatch(samples, la_nu, force, "mixed", ntime_monitor=200, tol=tol, rest_state=1)
        b.run(1, T_PROBE)
        probes = []
        for i, s in enumerate(samples):
            nz, ny, nx = s.walls.shape
            u, v, w, rho = (np.zeros((nz + 2, ny + 2, nx + 2)) for _ in range(4))
            b._call("download_macro", i, native.dptr(u), native.dptr(v), native.dptr(w), native.dptr(rho))
            probes.append(normalized_fields(u, v, w, rho, batch.global_walls(s) == 0, la_nu, force))
        r = b.run_steady(max_steps)
        for i, s in enumerate(samples):
            nz, ny, nx = s.walls.shape
            u, v, w, rho = (np.zeros((nz + 2, ny + 2, nx + 2)) for _ in range(4))
            b._call("download_macro", i, native.dptr(u), native.dptr(v), native.dptr(w), native.dptr(rho))
            fl = (batch.global_walls(s) == 0).astype(np.float32)
            j = done + i
            device_features(x[j], fl > 0, (True, bool(s.jper), False))
            x[j, ..., NFEAT:NIN] = torch.from_numpy(probes[i]).to("cuda", torch.bfloat16)
            tgt = normalized_fields(u, v, w, rho, fl > 0, la_nu, force)
            t[j, ..., :USED] = torch.from_numpy(tgt.astype(np.float32)).to("cuda", torch.bfloat16)
            sq += (tgt[fl > 0] ** 2).sum(axis=0)
            npos += int((fl > 0).sum())
            steps[j] = r.conv_step[i]
        b.close()
        done += m
        k += 1
        if log:
            log(f"dataset: {done}/{n} fields, {time.perf_counter() - t0:.1f}s")
    torch.cuda.synchronize()
    return Dataset(x, t, steps, time.perf_counter() - t0, np.sqrt(sq / max(npos, 1)))


def _mirror(x, t, axis: int, channel: int):
    """A sample mirrored across x (axis 3, channel 0 = u) or y (axis 2, channel 1 = v): the walls stay walls,
    the mirrored steady flow is the steady flow of the mirrored pack with that velocity component negated."""
    x, t = x.flip(axis), t.flip(axis)
    t[..., channel] = -t[..., channel]
    return x, t


def train(net: WarmStartNet, data: Dataset, epochs: int, batch_size: int = 4, lr: float = 1e-3,
          seed: int = 0, log=None, lr_final: Optional[float] = None, mirror: bool = True) -> List[float]:
    """Adam over shuffled mini-batches, the step size decaying from lr to lr_final (default lr / 10) on a
    cosine over the run; with ``mirror`` each mini-batch is mirrored across x and/or y at random; returns the
    mean loss of each epoch."""
    torch = net.torch
    n = data.x.shape[0]
    rng = np.random.default_rng(seed)
    lr_final = lr / 10.0 if lr_final is None else lr_final
    total = max(1, epochs * (n // batch_size))
    it = 0
    hist = []
    for e in range(epochs):
        order = rng.permutation(n)
        tot, cnt = 0.0, 0
        for s in range(0, n - batch_size + 1, batch_size):
            idx = torch.from_numpy(order[s:s + batch_size]).to("cuda")
            xb, tb = data.x.index_select(0, idx), data.t.index_select(0, idx)
            if mirror:
                if rng.random() < 0.5:
                    xb, tb = _mirror(xb, tb, 3, 0)
                if rng.random() < 0.5:
                    xb, tb = _mirror(xb, tb, 2, 1)
            loss = net.loss_and_backward(xb.contiguous(), tb.contiguous())
            net.adam(lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / total)))
            it += 1
            tot += loss
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {e}: loss {hist[-1]:.5g}")
    return hist


def probe_fields(sim, case, la_nu: float, force: float) -> np.ndarray:
    """Run ``sim`` (an SPSimulation started from rest at the collision's equilibrium) for T_PROBE steps and
    return its normalized fields (the network's probe input)."""
    sim.run(1, T_PROBE)
    m = sim.macro()
    return normalized_fields(m["u"], m["v"], m["w"], m["rho"], case.walls_global == 0, la_nu, force)


def warm_fields(net: WarmStartNet, case, la_nu: float, force: float, probe: Optional[np.ndarray] = None):
    """(rho, (u, v, w)) halo-1 arrays of ``case`` (an SPCase) from the network: u = prediction * F / nu,
    rho = 1 + P_SCALE F p with p shifted to zero mean over the fluid; ``probe`` the fields after T_PROBE steps
    of the same run (``probe_fields``)."""
    fl = (case.walls_global == 0)
    y = net.predict(fl.astype(np.float32), (bool(case.rp.kper), bool(case.rp.jper), False), probe).astype(np.float64)
    vel = [np.pad(y[m] * (force / la_nu), 1) for m in range(3)]
    p = y[3] - y[3][fl].mean()
    rho = np.pad(np.where(fl, 1.0 + P_SCALE * force * p, 1.0), 1, constant_values=1.0)
    return rho, vel


def warm_state(net: WarmStartNet, case, la_nu: float, force: float, probe: Optional[np.ndarray] = None):
    """Host-built initial state from the network's fields (populations at the collision's equilibrium); a
    running SPSimulation takes the same state on the device with ``sim.set_equilibrium(*warm_fields(...))``."""
    rho, vel = warm_fields(net, case, la_nu, force, probe)
    return case.collision_equilibrium_state(rho=rho, velocity=vel)


@dataclasses.dataclass
class RunRecord:
    steps: int
    seconds: float
    permeability: float
    status: str = "converged"   # or the run_steady status it stopped with ("max_steps", "failed")


@dataclasses.dataclass
class WarmStartResult:
    """One unseen sample run three ways to the same criterion: the plain path (the reference's initial state,
    ``plain_precision``), from rest at the collision's equilibrium and from the network's fields (both in
    ``precision``)."""
    side: int
    plain: RunRecord
    cold: RunRecord
    warm: RunRecord
    seconds_predict: float   # the probe steps, the prediction and the upload: part of warm.seconds

    @property
    def k_rel_err(self) -> float:
        """permeability of the warm-started run against the plain path's"""
        return abs(self.warm.permeability / self.plain.permeability - 1.0)

    @property
    def time_ratio(self) -> float:
        """wall time of the warm-started run (probe and prediction included) over the plain path's"""
        return self.warm.seconds / self.plain.seconds


def compare_runs(net: WarmStartNet, sample: batch.REVSample, la_nu: float = 0.1, force: float = 1e-6,
                 tol: float = 1e-7, precision: str = "fp32", plain_precision: str = "fp64",
                 max_steps: int = 600000, ntime_monitor: int = 500) -> WarmStartResult:
    """The same sample run to the reference's criterion from the reference's initial state, from rest at the
    collision's equilibrium and from the network's fields."""
    from .singlephase import SPCase, SPSimulation, permeability
    nx, ny, nz = sample.shape
    ctl = {"lattice_dimensions": (nx, ny, nz), "periodic_indicator": (0, sample.jper, 1),
           "domain_wall_status_y": (0, 0) if sample.jper else (1, 1), "domain_wall_status_z": (0, 0),
           "MPI_async_layers_num": (0, 4 if sample.jper else 0, 4), "fluid_viscosity": la_nu, "body_force_0": force,
           "MRT_collision_parameter_preset": 1, "steady_state_option": 1, "convergence_criteria": tol}
    runs = {}
    tp = 0.0
    for mode in ("plain", "cold", "warm"):
        case = SPCase(ctl, sample.walls, plain_precision if mode == "plain" else precision)
        st = case.initial_state() if mode == "plain" else case.collision_equilibrium_state()
        sim = SPSimulation(case, st)
        sim.synchronize()
        t0 = time.perf_counter()
        n0 = 1
        if mode == "warm":   # T_PROBE steps from rest, the network, then the run from its fields
            t1 = time.perf_counter()
            probe = probe_fields(sim, case, la_nu, force)
            sim.set_equilibrium(*warm_fields(net, case, la_nu, force, probe))
            n0 = T_PROBE + 1
            sim.synchronize()
            tp = time.perf_counter() - t1
        n, status, hist = sim.run_steady(n0, max_steps, ntime_monitor)
        dt = time.perf_counter() - t0
        sim.close()
        runs[mode] = RunRecord(n, dt, permeability(case, hist[-1][1]), status)
    return WarmStartResult(nx, runs["plain"], runs["cold"], runs["warm"], tp)
