"""MI355X engine for MoDLE's simulate hot path: the per-cell loop-extrusion epoch loop and the
banded contact matrix, behind the C-ABI of include/modle_gpu.h (libmodle_gpu.so)."""
from ._lib import MODE_DETERMINISTIC, MODE_FAST, LibraryMissing, MgpuError, lib  # noqa: F401

__all__ = ["MODE_DETERMINISTIC", "MODE_FAST", "LibraryMissing", "MgpuError", "lib"]
