"""Builds libmodle_gpu.so (the C-ABI library) in-tree with hipcc for gfx950.

    python -m modlefork_amd.build        # or __graft_entry__.build()

Every HIP translation unit is compiled with packed-FP32 instruction selection disabled; the
deterministic-mode kernels additionally with -ffp-contract=off (the reference performs no FMA
contraction, and bit-exact bands depend on it).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmodle_gpu.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("MGPU_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]
UNITS = [
    ("det_kernels.hip", ["-ffp-contract=off"]),
    ("fast_kernels.hip", []),
    ("mgpu_api.cpp", ["-ffp-contract=off"]),
    ("sg_kernels.hip", []),
    ("sg_model.cpp", []),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit):
    src, extra = unit
    obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
    cmd = [hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


HDF5_PREFIX = os.environ.get("MGPU_HDF5_PREFIX", "/opt/conda")
COOL_OUT = os.path.join(HERE, "libmodle_cool.so")


def build_cool(verbose: bool = False) -> str | None:
    """libmodle_cool.so, the .cool writer (host C++ on the HDF5 C library; include/modle_cool.h).
    Skipped, with a message, where the HDF5 headers are absent."""
    inc = os.path.join(HDF5_PREFIX, "include")
    lib = os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"HDF5 not found under {HDF5_PREFIX}: libmodle_cool.so not built", file=sys.stderr)
        return None
    tmp = COOL_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{inc}",
           os.path.join(CSRC, "cool_writer.cpp"), "-o", tmp, f"-L{lib}", "-lhdf5",
           f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for cool_writer.cpp:\n{r.stderr}")
    os.replace(tmp, COOL_OUT)
    if verbose:
        print(f"built {COOL_OUT}")
    return COOL_OUT


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_rng_tables.py"),
                    os.path.join(CSRC, "rng_tables.h")], check=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_jump_table.py"),
                    os.path.join(CSRC, "det_jump.h")], check=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_cool(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
