"""Synthetic wall arrays (1 = solid, 0 = fluid), indexed [k, j, i] with x fastest.

The reference ships its sample rocks in a git submodule that is not available offline; these generators
produce the seeded synthetic geometries the benchmarks are defined on:

* ``tube_sphere``: solid tube around the z axis plus a centred solid sphere, the shape the reference's
  built-in geometry uses (multiphase_3D/0.src/Misc.F90:213-244) with free radii.
* ``duct_obstacle``: the single-phase program's built-in obstacle, a solid sphere off the duct axis
  (singlephase_3D/0.src/Misc.F90:212-243).
* ``sphere_pack``: randomly placed overlapping solid spheres (radius uniform in [rmin, rmax]) added until
  the porosity of the sample reaches the target, with fluid buffer layers at both z ends.
"""
from __future__ import annotations

import numpy as np


def tube_sphere(nx: int, ny: int, nz: int, r_tube: float, r_sphere: float, buffer: int = 0) -> np.ndarray:
    """Tube wall where (i-xc)^2+(j-yc)^2 > r_tube^2 (outside the buffers) and a sphere where
    (i-xc)^2+(j-yc)^2+(k-zc)^2 < r_sphere^2; centre = 0.5*(n+1) in 1-based lattice coordinates."""
    xc, yc, zc = 0.5 * (nx + 1), 0.5 * (ny + 1), 0.5 * (nz + 1)
    k = np.arange(1, nz + 1, dtype=np.float64)[:, None, None]
    j = np.arange(1, ny + 1, dtype=np.float64)[None, :, None]
    i = np.arange(1, nx + 1, dtype=np.float64)[None, None, :]
    w = np.zeros((nz, ny, nx), dtype=np.int8)
    sphere = (i - xc) ** 2 + (j - yc) ** 2 + (k - zc) ** 2 < r_sphere ** 2
    kk = np.arange(1, nz + 1)[:, None, None]
    tube = ((i - xc) ** 2 + (j - yc) ** 2 > r_tube ** 2) & (kk > buffer) & (kk < nz - buffer + 1)
    w[np.broadcast_to(sphere, w.shape)] = 1
    w[np.broadcast_to(tube, w.shape)] = 1
    return w


def duct_obstacle(nx: int, ny: int, nz: int) -> np.ndarray:
    """Solid sphere of radius 0.15*ny centred at (0.5(nx+1), 0.45(ny+1), 0.3(nz+1)) in 1-based lattice
    coordinates, where (i-xc)^2+(j-yc)^2+(k-zc)^2 < r^2 (modify_geometry, singlephase_3D/0.src/Misc.F90:212-243)."""
    xc, yc, zc = 0.5 * (nx + 1), 0.45 * (ny + 1), 0.3 * (nz + 1)
    r1 = 0.15 * ny
    k = np.arange(1, nz + 1, dtype=np.float64)[:, None, None]
    j = np.arange(1, ny + 1, dtype=np.float64)[None, :, None]
    i = np.arange(1, nx + 1, dtype=np.float64)[None, None, :]
    return ((i - xc) ** 2 + (j - yc) ** 2 + (k - zc) ** 2 < r1 ** 2).astype(np.int8)


class SplitMix64:
    """SplitMix64 generator (Steele, Lea & Flood 2014); deterministic across platforms."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def sphere_pack(nx: int, ny: int, nz: int, porosity: float, rmin: float, rmax: float, seed: int,
                buffer: int = 8, walls_xy: bool = True) -> np.ndarray:
    """Overlapping solid spheres in the sample region (z layers buffer+1 .. nz-buffer, x/y interior when
    walls_xy) until the sample porosity drops to the target; buffers stay fluid.  Centres are uniform in
    the sample box; spheres are clipped to the sample region."""
    w = np.zeros((nz, ny, nx), dtype=np.int8)
    x0, x1 = (1, nx - 1) if walls_xy else (0, nx)
    y0, y1 = (1, ny - 1) if walls_xy else (0, ny)
    z0, z1 = buffer, nz - buffer
    sample = w[z0:z1, y0:y1, x0:x1]
    total = sample.size
    solid = 0
    target_solid = (1.0 - porosity) * total
    rng = SplitMix64(seed)
    sx, sy, sz = x1 - x0, y1 - y0, z1 - z0
    while solid < target_solid:
        cx = rng.uniform() * sx
        cy = rng.uniform() * sy
        cz = rng.uniform() * sz
        r = rmin + (rmax - rmin) * rng.uniform()
        ilo, ihi = max(int(np.floor(cx - r)), 0), min(int(np.ceil(cx + r)) + 1, sx)
        jlo, jhi = max(int(np.floor(cy - r)), 0), min(int(np.ceil(cy + r)) + 1, sy)
        klo, khi = max(int(np.floor(cz - r)), 0), min(int(np.ceil(cz + r)) + 1, sz)
        if ilo >= ihi or jlo >= jhi or klo >= khi:
            continue
        kk = (np.arange(klo, khi) + 0.5 - cz)[:, None, None]
        jj = (np.arange(jlo, jhi) + 0.5 - cy)[None, :, None]
        ii = (np.arange(ilo, ihi) + 0.5 - cx)[None, None, :]
        ball = (ii * ii + jj * jj + kk * kk) < r * r
        sub = sample[klo:khi, jlo:jhi, ilo:ihi]
        newly = ball & (sub == 0)
        solid += int(newly.sum())
        sub[newly] = 1
    return w


def porosity(walls: np.ndarray) -> float:
    return float((walls <= 0).sum()) / walls.size
