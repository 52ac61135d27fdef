"""MF-LBM per-time-step lattice update on AMD Instinct MI355X (gfx950).

Host-side mirror of the reference's driver for the hot path: the run inputs (``Case``), the GPU run
(``Simulation``) behind the C ABI of ``include/mflbm.h``, and the reference's on-disk formats (``io``).
"""
from . import geometry, io  # noqa: F401
from .case import BoundaryLists, Case, State  # noqa: F401
from .solver import MonitorResult, Simulation, post_monitor  # noqa: F401

__all__ = ["geometry", "io", "Case", "State", "BoundaryLists", "Simulation", "MonitorResult", "post_monitor"]
