"""Test-side binding of the CPU oracle (oracle/liboracle.so) behind the same driver class as the GPU path."""
from __future__ import annotations

import ctypes as C
import os

from mflbm import native
from mflbm.singlephase import SPSimulation
from mflbm.solver import Simulation

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")

_lib = None


def load_oracle():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so missing: build it with python -m mflbm.build")
        _lib = C.CDLL(ORACLE_LIB)
        _lib.oracle_set_threads.restype = C.c_int
        _lib.oracle_set_threads.argtypes = [C.c_int]
    return _lib


def set_threads(n: int) -> int:
    return load_oracle().oracle_set_threads(n)


class OracleSimulation(Simulation):
    prefix = "oracle_"

    def _load(self):
        return load_oracle()


class SPOracleSimulation(SPSimulation):
    """Single-phase oracle (oracle_sp_* entry points) behind the GPU driver class."""
    prefix = "oracle_"

    def _load(self):
        return load_oracle()


_ = native  # the oracle shares the ABI's struct layouts
