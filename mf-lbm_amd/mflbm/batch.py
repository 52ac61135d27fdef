"""Batched single-phase permeability runs: many REV subvolumes on one GPU (``mflbm_spb_*`` C ABI).

The reference computes one sample per job (singlephase_3D/0.src/Main.F90:124-230) and leaves the permeability
to its users.  Here every sample is the reference's single-phase problem (solid x walls, z periodic, y walls or
y periodic, body force along z), all samples advance in one lattice-update launch per step, and each stops at
the reference's flow-rate criterion (Monitor.F90:145-156), checked on the device every ``ntime_monitor`` steps.

* ``REVBatch(samples, ...)``: the device batch.  ``run_steady`` returns per-sample convergence steps, flow
  rates and permeabilities k = nu * Q / (A_xy * F) with the reference's A_xy (Initialization.F90:40-43).
* ``rev_samples(n, seed)``: the synthetic REV set: cubes of side drawn from {48, 64, 96, 128}, each a seeded
  sphere pack of porosity drawn from [0.15, 0.35] with 4 fluid layers at each z end.
* ``autotune(samples, ...)``: times the lattice update for each (tile rows, block size) candidate on a
  representative subset of the samples and returns the fastest; results are cached on disk by batch signature.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import hashlib
import json
import os
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import geometry, io, native
from .native import i8ptr

PRECISION = {"fp64": 0, "fp32": 1, "mixed": 2}


@dataclasses.dataclass
class REVSample:
    walls: np.ndarray      # (nz, ny, nx) int8, 1 = solid
    jper: int = 0          # y periodic (else solid y walls)

    @property
    def shape(self) -> Tuple[int, int, int]:
        nz, ny, nx = self.walls.shape
        return nx, ny, nz


def _f32(x: float) -> float:
    return float(np.float32(x))


def area_xy(nx: int, ny: int) -> float:
    """A_xy of the single-phase program: la_x * la_y with la = n - 1 - 0.5 - 0.5, evaluated in single precision
    as the reference does (Initialization.F90:40-43, whatever the y boundary)."""
    la_x = _f32(_f32(_f32(nx - 1) - 0.5) - 0.5)
    la_y = _f32(_f32(_f32(ny - 1) - 0.5) - 0.5)
    return la_x * la_y


def global_walls(sample: REVSample) -> np.ndarray:
    """walls_global of the sample's run (set_walls, Misc.F90:18-109): solid x walls at i = 1 and nx, solid y
    walls at j = 1 and ny unless y is periodic, z periodic."""
    nx, ny, nz = sample.shape
    wy = (0, 0) if sample.jper else (1, 1)
    return io.global_walls(sample.walls, nx, ny, nz, (1, 1), wy, (0, 0))


def walls_h2(sample: REVSample) -> np.ndarray:
    """The sample's walls with halo 2 (set_walls through mflbm_setup_walls): x walls, z periodic."""
    lib = native.load()
    nx, ny, nz = sample.shape
    wp = native.Params()
    wp.nx, wp.ny, wp.nz, wp.jper, wp.kper = nx, ny, nz, sample.jper, 1
    wp.wall_x_min = wp.wall_x_max = 1
    wp.wall_y_min = wp.wall_y_max = 0 if sample.jper else 1
    wp.wall_z_min = wp.wall_z_max = 0
    out = np.zeros((nz + 4, ny + 4, nx + 4), dtype=np.int8)
    wg = np.ascontiguousarray(global_walls(sample), dtype=np.int8)
    rc = lib.mflbm_setup_walls(C.byref(wp), i8ptr(wg), i8ptr(out))
    if rc != 0:
        raise RuntimeError(f"mflbm_setup_walls failed with code {rc}")
    return out


def rev_samples(n: int, seed: int = 2024, sides: Sequence[int] = (48, 64, 96, 128),
                porosity: Tuple[float, float] = (0.15, 0.35), radius: Tuple[float, float] = (4.0, 8.0),
                buffer: int = 4) -> List[REVSample]:
    """n REV cubes: side and porosity drawn per sample from a SplitMix64 stream of ``seed``; each cube is a
    sphere pack (geometry.sphere_pack, solid x/y walls) seeded by the stream."""
    rng = geometry.SplitMix64(seed)
    out = []
    for _ in range(n):
        side = sides[int(rng.uniform() * len(sides))]
        phi = porosity[0] + (porosity[1] - porosity[0]) * rng.uniform()
        s = int(rng.next_u64() & 0x7FFFFFFF)
        out.append(REVSample(geometry.sphere_pack(side, side, side, phi, radius[0], radius[1], s, buffer=buffer)))
    return out


@dataclasses.dataclass
class BatchResult:
    conv_step: np.ndarray      # per sample: step of convergence (> 0), failure (< 0), 0 still running
    flowrate: np.ndarray       # per sample: last flow rate Q = sum(w) / nz
    permeability: np.ndarray   # nu * Q / (A_xy * F)
    periods: int
    seconds: float


class REVBatch:
    """The samples of one batch on the GPU."""

    def __init__(self, samples: Sequence[REVSample], la_nu: float = 0.1, force_z: float = 1e-6,
                 precision: str = "fp32", mrt_preset: int = 1, ntime_monitor: int = 500, tol: float = 1e-7,
                 tile_rows: int = 0, block: int = 0, rest_state: int = 0):
        """rest_state 1 starts every sample at the collision's rest equilibrium instead of the lattice weights
        (SPCase.collision_equilibrium_state)."""
        if precision not in PRECISION:
            raise ValueError(f"precision must be one of {sorted(PRECISION)}")
        self.lib = native.load()
        self.fn = native.bind(self.lib, "mflbm_", batch=True)
        self.samples = list(samples)
        self.la_nu, self.force_z = la_nu, force_z
        p = native.SPBParams()
        p.precision, p.mrt_preset, p.ntime_monitor = PRECISION[precision], mrt_preset, ntime_monitor
        p.tile_rows, p.block, p.rest_state = tile_rows, block, rest_state
        p.la_nu, p.force_z, p.tol = la_nu, force_z, tol
        self.params = p
        self.ctx = C.c_void_p()
        rc = self.fn["create"](C.byref(p), C.byref(self.ctx))
        if rc != 0:
            raise RuntimeError(f"mflbm_spb_create failed ({rc}): {self.lib.mflbm_create_error().decode()}")
        for s in self.samples:
            nx, ny, nz = s.shape
            self._call("add_sample", nx, ny, nz, s.jper, i8ptr(walls_h2(s)))
        self._call("commit")
        ns, nfl, nb = C.c_int32(), C.c_int64(), C.c_int64()
        self._call("info", C.byref(ns), C.byref(nfl), C.byref(nb))
        self.fluid_nodes = nfl.value
        self.pop_bytes = nb.value
        self.area = np.array([area_xy(*s.shape[:2]) for s in self.samples])

    def _call(self, name: str, *args) -> None:
        rc = self.fn[name](self.ctx, *args)
        if rc != 0:
            m = self.fn["last_error"](self.ctx)
            raise RuntimeError(f"mflbm_spb_{name} failed ({rc}): {m.decode() if m else ''}")

    def close(self) -> None:
        if self.ctx:
            self.fn["destroy"](self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sample_fluid(self, i: int) -> int:
        v = C.c_int64()
        self._call("sample_fluid", i, C.byref(v))
        return v.value

    def run(self, ntime0: int, nsteps: int) -> None:
        self._call("run", ntime0, nsteps)

    def synchronize(self) -> None:
        self._call("synchronize")

    def state(self, i: int) -> np.ndarray:
        nx, ny, nz = self.samples[i].shape
        f = np.zeros((19, nz + 2, ny + 2, nx + 2))
        self._call("download_state", i, native.ptr_array(list(f)))
        return f

    def macro(self, i: int) -> Dict[str, np.ndarray]:
        """u, v, w, rho of sample i (after an even step) in the halo-1 layout, solid nodes 0."""
        nx, ny, nz = self.samples[i].shape
        out = {k: np.zeros((nz + 2, ny + 2, nx + 2)) for k in ("u", "v", "w", "rho")}
        self._call("download_macro", i, *(native.dptr(out[k]) for k in ("u", "v", "w", "rho")))
        return out

    def upload(self, i: int, f: np.ndarray) -> None:
        f = np.ascontiguousarray(f, dtype=np.float64)
        self._keep = f
        self._call("upload_state", i, native.ptr_array(list(f)))

    def results(self, seconds: float = 0.0) -> BatchResult:
        n = len(self.samples)
        cs = np.zeros(n, np.int32)
        q = np.zeros(n)
        per = np.zeros(1, np.int32)
        self._call("results", native.iptr(cs), native.dptr(q), native.iptr(per))
        k = self.la_nu * q / (self.area * self.force_z)
        return BatchResult(cs, q, k, int(per[0]), seconds)

    def run_steady(self, max_steps: int, log=None, chunk_steps: int = 10000) -> BatchResult:
        """Monitor periods until every sample has converged (or failed) or max_steps steps have run; the
        periods go to the device in chunks of about chunk_steps steps, after each of which ``log`` (if given)
        receives a progress line."""
        nm = self.params.ntime_monitor
        left = C.c_int32()
        t0 = time.perf_counter()
        total = max(1, -(-max_steps // nm))
        chunk = max(1, chunk_steps // nm)
        done = 0
        while done < total:
            k = min(chunk, total - done)
            self._call("run_steady", k, C.byref(left))
            done += k
            if log:
                log(f"batch: {done * nm} steps, {left.value} samples running, {time.perf_counter() - t0:.1f}s")
            if left.value == 0:
                break
        return self.results(time.perf_counter() - t0)

    def reset_monitor(self) -> None:
        self._call("reset_monitor")


def _signature(samples: Sequence[REVSample], precision: str) -> str:
    shapes = sorted((s.shape, s.jper, int((s.walls == 0).sum()) // 4096) for s in samples)
    return hashlib.sha1(json.dumps([precision, shapes]).encode()).hexdigest()[:16]


def _subset(samples: Sequence[REVSample], per_class: int) -> List[REVSample]:
    """Up to per_class samples of each distinct shape, in order."""
    seen: Dict[tuple, int] = {}
    out = []
    for s in samples:
        if seen.get(s.shape, 0) < per_class:
            seen[s.shape] = seen.get(s.shape, 0) + 1
            out.append(s)
    return out


def autotune(samples: Sequence[REVSample], precision: str = "fp32", tiles: Sequence[int] = (8, 16, 32, 64),
             blocks: Sequence[int] = (128, 256, 512), steps: int = 40, per_class: int = 2,
             cache: Optional[str] = None, la_nu: float = 0.1, force_z: float = 1e-6) -> dict:
    """Pick (tile_rows, block) for a batch: time ``steps`` steps of the lattice update for every candidate on
    up to ``per_class`` samples of each shape, and return the fastest with all timings.  A cache file (default
    $MFLBM_AUTOTUNE_CACHE, else none) keyed by the batch signature skips the search next time."""
    cache = cache if cache is not None else os.environ.get("MFLBM_AUTOTUNE_CACHE")
    key = _signature(samples, precision)
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            db = json.load(fh)
        if key in db:
            return dict(db[key], cached=True)
    sub = _subset(samples, per_class)
    timings = []
    for ty in tiles:
        for nt in blocks:
            b = REVBatch(sub, la_nu, force_z, precision, tile_rows=ty, block=nt)
            b.run(1, 4)                 # warm-up
            b.synchronize()
            t0 = time.perf_counter()
            b.run(5, steps)
            b.synchronize()
            dt = (time.perf_counter() - t0) / steps
            timings.append({"tile_rows": ty, "block": nt, "ms_per_step": dt * 1e3,
                            "mlups": b.fluid_nodes / dt / 1e6})
            b.close()
    best = min(timings, key=lambda r: r["ms_per_step"])
    out = {"tile_rows": best["tile_rows"], "block": best["block"], "timings": timings,
           "subset_samples": len(sub), "steps": steps, "cached": False}
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                db = json.load(fh)
        db[key] = {k: v for k, v in out.items() if k != "cached"}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as fh:
            json.dump(db, fh, indent=1)
    return out
