"""Field output files of the reference, byte for byte: legacy VTK (structured points, binary, big-endian) and
the raw per-rank dumps used for very large runs.

Two-phase program (multiphase_3D/0.src/IO_multiphase.F90):
* ``write_full_flow`` (VTK_legacy_writer_3D type 1, :824-1003): phi, density, velocity_X/Y/Z in FP64 or FP32
  (``output_fieldData_precision_cmd``); type 3 (``write_force_vectors``) writes cn_x, cn_y, cn_z, c_norm and
  phi under the same five labels, as the reference does.
* ``write_phase`` (type 2): phi in FP32 with solid nodes written as 0.
* ``write_walls`` (VTK_walls_bin, :1104-1149): walls_global as 4-byte integers.
* ``save_phi`` / ``save_macro`` (:650-714): raw stream files with a 6-integer header.

Single-phase program (singlephase_3D/0.src/IO.F90): ``write_sp_full_flow`` (VTK_legacy_writer_3D, :591-714),
``write_sp_full_flow_half`` (:716-808, every second node, FP32), ``save_sp_macro`` (:550-585).

Arrays are indexed [k, j, i] (x fastest), interior only (nz, ny, nx), as ``Simulation.macro()`` returns them
after dropping the halo; file names follow the reference (``vtk_name``).
"""
from __future__ import annotations

from typing import Dict, Iterable, List, Tuple

import numpy as np

_LF = "\n"


def vtk_name(nt: int) -> str:
    """The reference's '(i10.10,".vtk")' file-name suffix."""
    return f"{nt:010d}.vtk"


def _i10(v: int) -> str:
    return f"{v:10d}"


def _header(nx: int, ny: int, nz: int, npoints: int) -> bytes:
    # each header line is an 80-character Fortran buffer written with trim(): the line feed ends it
    lines = ["# vtk DataFile Version 3.0", "vtk output", "BINARY", "DATASET STRUCTURED_POINTS ",
             "DIMENSIONS " + _i10(nx) + " " + _i10(ny) + " " + _i10(nz),
             "ORIGIN " + _i10(1) + " " + _i10(1) + " " + _i10(1),
             "SPACING " + _i10(1) + " " + _i10(1) + " " + _i10(1),
             "POINT_DATA " + f"{npoints:14d}"]
    return "".join(s + _LF for s in lines).encode("ascii")


_BE = {"float": ">f4", "double": ">f8", "int": ">i4"}


def _scalars(name: str, fmt: str, a: np.ndarray) -> bytes:
    # the field writers hold the type name in a character*30 variable, so it reaches the file blank-padded
    # to 30 characters; the walls writer spells 'int' as a literal
    t = fmt if fmt == "int" else f"{fmt:<30s}"
    head = ("SCALARS " + name + " " + t + _LF + "LOOKUP_TABLE default" + _LF).encode("ascii")
    return head + np.ascontiguousarray(a, dtype=_BE[fmt]).tobytes()


def _fmt(double: bool) -> str:
    return "double" if double else "float"


def _as32(a: np.ndarray) -> np.ndarray:
    # Fortran real(x) of a real(8): round to nearest single
    return np.asarray(a, dtype=np.float64).astype(np.float32)


def _dims(a: np.ndarray) -> Tuple[int, int, int]:
    nz, ny, nx = a.shape
    return nx, ny, nz


def _write(path: str, dims: Tuple[int, int, int], npoints: int, items: Iterable[Tuple[str, str, np.ndarray]]) -> None:
    with open(path, "wb") as fh:
        fh.write(_header(*dims, npoints))
        for name, fmt, a in items:
            fh.write(_scalars(name, fmt, a))


def write_full_flow(path: str, phi, rho, u, v, w, double: bool = True) -> None:
    """VTK_legacy_writer_3D type 1 of the two-phase program."""
    fmt = _fmt(double)
    conv = (lambda a: np.asarray(a, dtype=np.float64)) if double else _as32
    nx, ny, nz = _dims(phi)
    _write(path, (nx, ny, nz), nx * ny * nz,
           [("phi", fmt, conv(phi)), ("density", fmt, conv(rho)), ("velocity_X", fmt, conv(u)),
            ("velocity_Y", fmt, conv(v)), ("velocity_Z", fmt, conv(w))])


def write_force_vectors(path: str, cn_x, cn_y, cn_z, c_norm, phi, double: bool = True) -> None:
    """VTK_legacy_writer_3D type 3: the normal field under the labels of type 1 (phi <- phi, density <- c_norm,
    velocity_X/Y/Z <- cn_x/y/z), exactly as the reference fills its gather buffers."""
    write_full_flow(path, phi, c_norm, cn_x, cn_y, cn_z, double)


def write_phase(path: str, phi, walls_global) -> None:
    """VTK_legacy_writer_3D type 2: phi in single precision, solid nodes written as 0."""
    phi = np.where(np.asarray(walls_global) > 0, 0.0, np.asarray(phi, dtype=np.float64))
    nx, ny, nz = _dims(phi)
    _write(path, (nx, ny, nz), nx * ny * nz, [("phi", "float", _as32(phi))])


def write_walls(path: str, walls_global) -> None:
    """VTK_walls_bin: walls_global as 4-byte integers."""
    w = np.asarray(walls_global)
    nx, ny, nz = _dims(w)
    _write(path, (nx, ny, nz), nx * ny * nz, [("walls", "int", w.astype(np.int32))])


def write_sp_full_flow(path: str, rho, u, v, w, double: bool = True) -> None:
    """Single-phase VTK_legacy_writer_3D: density, velocity_X/Y/Z."""
    fmt = _fmt(double)
    conv = (lambda a: np.asarray(a, dtype=np.float64)) if double else _as32
    nx, ny, nz = _dims(rho)
    _write(path, (nx, ny, nz), nx * ny * nz,
           [("density", fmt, conv(rho)), ("velocity_X", fmt, conv(u)), ("velocity_Y", fmt, conv(v)),
            ("velocity_Z", fmt, conv(w))])


def write_sp_full_flow_half(path: str, rho, u, v, w) -> None:
    """Single-phase VTK_legacy_writer_3D_half: every second node from the first, FP32; the header says
    n/2 per dimension and nx*ny*nz/8 points (integer division), as the reference writes it."""
    nx, ny, nz = _dims(rho)
    sub = (slice(0, None, 2),) * 3
    _write(path, (nx // 2, ny // 2, nz // 2), nx * ny * nz // 8,
           [("density", "float", _as32(np.asarray(rho)[sub])), ("velocity_X", "float", _as32(np.asarray(u)[sub])),
            ("velocity_Y", "float", _as32(np.asarray(v)[sub])), ("velocity_Z", "float", _as32(np.asarray(w)[sub]))])


def _raw(path: str, arrays: List[np.ndarray], double: bool, idx=(0, 0, 0)) -> None:
    nx, ny, nz = _dims(arrays[0])
    with open(path, "wb") as fh:
        fh.write(np.array(list(idx) + [nx, ny, nz], dtype="<i4").tobytes())
        for a in arrays:
            fh.write(np.ascontiguousarray(a, dtype="<f8" if double else "<f4").tobytes())


def save_phi(path: str, phi, double: bool = True, idx=(0, 0, 0)) -> None:
    """save_phi: idx, idy, idz, nx, ny, nz (int32, native little-endian) then phi."""
    _raw(path, [np.asarray(phi, dtype=np.float64) if double else _as32(phi)], double, idx)


def save_macro(path: str, u, v, w, rho, phi, double: bool = True, idx=(0, 0, 0)) -> None:
    """save_macro of the two-phase program: header, then u, v, w, rho, phi."""
    arrs = [u, v, w, rho, phi]
    _raw(path, [np.asarray(a, dtype=np.float64) if double else _as32(a) for a in arrs], double, idx)


def save_sp_macro(path: str, u, v, w, rho, double: bool = True, idx=(0, 0, 0)) -> None:
    """save_macro of the single-phase program: header, then u, v, w, rho."""
    arrs = [u, v, w, rho]
    _raw(path, [np.asarray(a, dtype=np.float64) if double else _as32(a) for a in arrs], double, idx)


def read_vtk(path: str) -> Dict[str, object]:
    """Parse a file written by the writers above (or by the reference): dimensions, point count and the
    named scalar arrays, shaped [k, j, i]."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0

    def line() -> str:
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode("ascii")
        pos = e + 1
        return s

    head = [line() for _ in range(8)]
    dims = tuple(int(x) for x in head[4].split()[1:4])
    npts = int(head[7].split()[1])
    out: Dict[str, object] = {"dims": dims, "points": npts, "header": head}
    nx, ny, nz = dims
    n = nx * ny * nz
    while pos < len(data):
        name, fmt = line().split()[1:3]   # split() drops the blank padding
        line()   # LOOKUP_TABLE default
        dt = np.dtype(_BE[fmt])
        a = np.frombuffer(data, dtype=dt, count=n, offset=pos)
        pos += n * dt.itemsize
        out[name] = a.reshape(nz, ny, nx)
    return out
