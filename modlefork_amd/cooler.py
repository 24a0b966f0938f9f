"""ctypes binding of libmodle_cool.so (include/modle_cool.h): .cool output of the simulate path.

`write_cool(path, results, bin_size)` plays the role of Simulation::write_contacts_to_disk
(src/libmodle/cpu/simulation.cpp:113-210) with its Cooler<u32> writer: one call per chromosome in
genome order, the band of a simulated chromosome or, for a chromosome that was not simulated,
its bins and index only. `results` maps chromosome names to (band, missed, plan) as
Simulation.run_simulate returns them.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmodle_cool.so")

_lib = None


class CoolError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CoolError(f"{LIB_PATH} is missing: build it with modlefork_amd.build")
        L = C.CDLL(LIB_PATH)
        L.mcool_open.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.c_char_p,
                                 C.POINTER(C.c_void_p)]
        L.mcool_write_metadata.argtypes = [C.c_void_p, C.c_char_p]
        L.mcool_write_chrom.argtypes = [C.c_void_p, C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                        C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64, C.c_uint64]
        L.mcool_close.argtypes = [C.c_void_p]
        L.mcool_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CoolError(f"libmodle_cool error {rc}: {lib().mcool_last_error().decode()}")


class CoolerWriter:
    """Cooler<u32> opened for writing (cooler_impl.hpp:22-93)."""

    def __init__(self, path: str, bin_size: int, max_name_length: int, assembly: str = ""):
        self.w = C.c_void_p()
        _check(lib().mcool_open(path.encode(), bin_size, max_name_length, assembly.encode(),
                                C.byref(self.w)))

    def write_metadata(self, meta: dict | str):
        text = meta if isinstance(meta, str) else json.dumps(meta)
        _check(lib().mcool_write_metadata(self.w, text.encode()))

    def write_chrom(self, name: str, start: int, end: int, length: int, band=None,
                    nrows: int = 0, ncols: int = 0):
        b = name.encode()
        ptr = None
        if band is not None:
            band = np.ascontiguousarray(band, np.uint32)
            assert len(band) == nrows * ncols + 1
            ptr = band.ctypes.data_as(C.POINTER(C.c_uint32))
        _check(lib().mcool_write_chrom(self.w, b, len(b), start, end, length, ptr, nrows, ncols))

    def close(self):
        if self.w:
            w, self.w = self.w, C.c_void_p()
            _check(lib().mcool_close(w))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_cool(path: str, chroms, results: dict, bin_size: int, metadata: dict | None = None,
               assembly: str = ""):
    """chroms: [(name, size)] in genome order; results: {name: (band, missed, plan)} for the
    simulated chromosomes (whole chromosomes, [0, size)). Chromosomes without a result get their
    bins and index only, as the reference does for chromosomes it did not simulate."""
    width = max(len(n.encode()) for n, _ in chroms)
    with CoolerWriter(path, bin_size, width, assembly) as w:
        if metadata:
            w.write_metadata(metadata)
        for name, size in chroms:
            if name in results:
                band, _, plan = results[name]
                w.write_chrom(name, 0, size, size, band, plan.nrows, plan.ncols)
            else:
                w.write_chrom(name, 0, size, size)
