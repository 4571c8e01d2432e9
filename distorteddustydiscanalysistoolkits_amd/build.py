"""Build libpartia_mi355x.so in-tree with hipcc for gfx950.

    python -m distorteddustydiscanalysistoolkits_amd.build

Every translation unit is compiled with -ffp-contract=off (FP64 kernels reproduce the
reference's un-fused arithmetic; FP32 kernels write fma() explicitly where they want it)
and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpartia_mi355x.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["pmx_api.hip", "pmx_build.hip", "pmx_interp64.hip", "pmx_interp32.hip", "pmx_lines_tiled.hip", "pmx_batch.hip", "pmx_phantom.hip",
           "pmx_noalloc.hip", "pmx_lut.cpp"]
ARCH = os.environ.get("PMX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags():
    # -Werror=shadow: a kernel-local name that hides an outer one captured by a lambda (a tile
    # bound, a counter) silently changes which value the lambda reads
    return ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
            "-Wno-unused-variable", "-Werror=shadow", "-I" + os.path.join(HERE, "..", "include")]


def compile_one(src: str) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags()
    if src.endswith(".hip"):
        cmd += ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


# The GridDataset writer / reader (pmx_griddataset.cpp): a host library of its own on the HDF5 C
# library, so that the interpolation library does not depend on HDF5.  libstdc++ is linked
# statically: the HDF5 directory's own C++ runtime must not be picked up through the run path.
GD_OUT = os.path.join(HERE, "libpartia_griddataset.so")


def hdf5_dir():
    for cand in (os.environ.get("PMX_HDF5_DIR"), "/opt/conda", "/usr"):
        if cand and os.path.exists(os.path.join(cand, "include", "hdf5.h")):
            return cand
    return None


def build_griddataset(verbose: bool = False):
    h5 = hdf5_dir()
    if h5 is None:
        if verbose:
            print("HDF5 headers not found: libpartia_griddataset.so not built")
        return None
    libdir = os.path.join(h5, "lib")
    tmp = GD_OUT + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-Werror=shadow",
           "-I" + os.path.join(h5, "include"), "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(CSRC, "pmx_griddataset.cpp"), "-o", tmp, "-L" + libdir, "-lhdf5",
           "-Wl,-rpath," + libdir, "-static-libstdc++", "-static-libgcc", "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, GD_OUT)
    if verbose:
        print("built", GD_OUT)
    return GD_OUT


def build(verbose: bool = False) -> str:
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_griddataset(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
