#!/usr/bin/env python3
"""Boundary-info files written by the reference's preprocessing tool (oracle/_ref/wall_boundary_process, built
by oracle/build_ref.sh from preprocessing/3.wall_boundary_preprocess/3d_wall_boundary_process.f90) for the
walls of fixture cases.  Output: tests/golden/boundary_<case>.npz holding the file bytes ("data") and the tool's
input flags ("flags": periodic x/y/z, walls x/y/z min/max).  The tool is used only through the built binary.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(REPO, "mf-lbm_amd"))

from mflbm import io  # noqa: E402

TOOL = os.path.join(REPO, "oracle", "_ref", "wall_boundary_process")
CASES = ["open_velocity", "periodic_tube"]


def main() -> None:
    if not os.path.exists(TOOL):
        sys.exit("tool missing: run oracle/build_ref.sh first")
    for name in CASES:
        z = np.load(os.path.join(HERE, name + ".npz"))
        ctl = json.loads(str(z["control"]))
        per = ctl.get("periodic_indicator", [0, 0, 1])
        wx = ctl.get("domain_wall_status_x", [1, 1])
        wy = ctl.get("domain_wall_status_y", [1, 1])
        wz = ctl.get("domain_wall_status_z", [0, 0])
        with tempfile.TemporaryDirectory() as wd:
            os.makedirs(os.path.join(wd, "output"))
            io.write_walls(os.path.join(wd, "walls.dat"), z["walls"])
            with open(os.path.join(wd, "input_parameters.txt"), "w") as fh:
                fh.write("wall_file_path:\n" + os.path.join(wd, "walls.dat") + "\n")
                fh.write(f"{per[0]},{per[1]},{per[2]}\n{wx[0]},{wx[1]}\n{wy[0]},{wy[1]}\n{wz[0]},{wz[1]}\n")
            subprocess.run([TOOL], cwd=wd, check=True, stdout=subprocess.DEVNULL)
            with open(os.path.join(wd, "output", "indirect_addressing_walls_boundary.dat"), "rb") as fh:
                data = np.frombuffer(fh.read(), dtype=np.uint8)
        np.savez_compressed(os.path.join(HERE, f"boundary_{name}.npz"), data=data,
                            flags=np.array(list(per) + list(wx) + list(wy) + list(wz), dtype=np.int32))
        print(name, data.size, "bytes")


if __name__ == "__main__":
    main()
