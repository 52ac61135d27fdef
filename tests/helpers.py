"""Shared test utilities: fixture loading, digests, state comparison."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from mflbm import Case
from mflbm.singlephase import SPCase

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL_CASES = ["periodic_tube", "open_velocity", "open_pressure", "porous_plate", "periodic_yz_drop"]
C1_CASE = "c1_tube_sphere_48x48x128"
SP_CASES = ["sp_periodic_obstacle", "sp_open_velocity", "sp_open_pressure", "sp_periodic_yz"]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def load_fixture(name: str):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    ctl = json.loads(str(z["control"]))
    ctl = {k: tuple(v) if isinstance(v, list) else v for k, v in ctl.items()}
    return z, ctl


def fixture_case(name: str):
    z, ctl = load_fixture(name)
    return z, Case(ctl, z["walls"])


def fixture_steps(z) -> list:
    return sorted(int(k.split("_")[1]) for k in z.files if k.startswith("sha_") and k.endswith("_f"))


def phi_as_reference_keeps_it(case: Case, phi: np.ndarray) -> np.ndarray:
    """Interior phi as the reference's checkpoint holds it.  The reference's start-up VTK writer zeroes phi
    at solid interior nodes (IO_multiphase.F90:877-887); the colour-gradient chain re-extrapolates the
    solid boundary nodes every step, the other solid nodes keep the 0.  The dynamics never read those
    nodes, so the emulation is applied here, to the compared copy only."""
    out = phi[4:-4, 4:-4, 4:-4].copy()
    rp = case.rp
    solid = case.walls_global == 1
    keep = np.zeros_like(solid)
    s = case.lists.s_ijk
    inside = (s[:, 0] >= 1) & (s[:, 0] <= rp.nx) & (s[:, 1] >= 1) & (s[:, 1] <= rp.ny) & (s[:, 2] >= 1) & (s[:, 2] <= rp.nz)
    si = s[inside]
    keep[si[:, 2] - 1, si[:, 1] - 1, si[:, 0] - 1] = True
    out[solid & ~keep] = 0.0
    return out


def state_digests(case: Case, st) -> dict:
    d = {"f": sha(st.f), "g": sha(st.g), "f_int": sha(st.f[:, 1:-1, 1:-1, 1:-1]),
         "g_int": sha(st.g[:, 1:-1, 1:-1, 1:-1]), "phi_int": sha(phi_as_reference_keeps_it(case, st.phi))}
    return d


def assert_matches_reference(z, case: Case, st, n: int) -> None:
    d = state_digests(case, st)
    for key in ("f", "g", "f_int", "g_int"):
        assert d[key] == str(z[f"sha_{n}_{key}"]), f"step {n}: {key} differs from the reference checkpoint"
    if n > 0:
        assert d["phi_int"] == str(z[f"sha_{n}_phi_int"]), f"step {n}: interior phi differs from the reference"


def max_rel_diff(a: np.ndarray, b: np.ndarray) -> float:
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def load_sp(name: str, precision: str = "fp64"):
    """Single-phase fixture and its case."""
    z, ctl = load_fixture(name)
    return z, SPCase(ctl, z["walls"], precision)


def assert_sp_matches(z, case: SPCase, st, n: int) -> None:
    assert sha(st.f) == str(z[f"sha_{n}_f"]), f"step {n}: populations differ from the reference checkpoint"
    if case.rp.outlet_bc == 1:
        assert sha(st.f_convec) == str(z[f"sha_{n}_convec"]), f"step {n}: convective history differs"
