#!/usr/bin/env python3
"""Golden fixtures for the field-output files: run the reference binaries built by oracle/build_ref.sh on one
two-phase and one single-phase case with VTK output switched on, and keep the sha256 digest, size and header
lines of every file they write (walls, phase-field, full-flow-field VTK, in FP64 and FP32).

tests/test_vtk.py writes the same files from the oracle's state with mflbm.vtk and compares digests, which
pins the writers byte for byte and the macroscopic fields (compute_macro_vars) bit for bit.
Output: tests/golden/vtk_files.npz.  Needs the reference binaries; never run on a machine without them.
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import make_fixtures as mp      # noqa: E402
import make_fixtures_sp as sp   # noqa: E402

STEPS = 10


def digest_tree(root: str) -> dict:
    out = {}
    for dirpath, _, files in os.walk(os.path.join(root, "out3.field_data")):
        for f in sorted(files):
            p = os.path.join(dirpath, f)
            with open(p, "rb") as fh:
                data = fh.read()
            rel = os.path.relpath(p, os.path.join(root, "out3.field_data"))
            head = data[:400].split(b"\n")[:8]
            out[rel] = {"sha256": hashlib.sha256(data).hexdigest(), "size": len(data),
                        "header": [h.decode("ascii", "replace") for h in head]}
    return out


def main() -> None:
    root = tempfile.mkdtemp(prefix="mflbm_vtk_")
    res = {}
    try:
        spec = mp.CASES["periodic_tube"]
        for prec in (1, 0):
            ctl = dict(spec["control"], ntime_visual=STEPS, ntime_animation=STEPS,
                       output_fieldData_precision_cmd=prec)
            wd = os.path.join(root, f"mp{prec}")
            mp.run_ref(wd, spec["geom"](), ctl, STEPS, 1)
            res[f"periodic_tube_prec{prec}"] = digest_tree(wd)
        spec = sp.CASES["sp_periodic_obstacle"]
        for prec in (1, 0):
            ctl = dict(spec["control"], ntime_visual=STEPS, output_fieldData_precision_cmd=prec)
            wd = os.path.join(root, f"sp{prec}")
            sp.run_ref(wd, spec["geom"](), ctl, STEPS, 1)
            res[f"sp_periodic_obstacle_prec{prec}"] = digest_tree(wd)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    np.savez_compressed(os.path.join(HERE, "vtk_files.npz"), files=json.dumps(res, indent=1, sort_keys=True),
                        steps=STEPS)
    print(json.dumps(res, indent=1, sort_keys=True)[:3000])


if __name__ == "__main__":
    main()
