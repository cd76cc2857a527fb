"""Compiles the gfx950 HIP library libdnsbox_hip.so in-tree.

hipcc cross-compiles for gfx950 without a GPU.  Each translation unit is
compiled in parallel, then linked into dnsbox_amd/lib/libdnsbox_hip.so, which
tests, smoke() and bench.py load through ctypes.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libdnsbox_hip.so")
OBJDIR = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["box_f64.hip", "box_f32.hip", "box_grid_f64.hip", "box_grid_f32.hip", "box_grid_f32s64.hip", "group.hip",
           "dnsbox_capi.hip", "noalloc.hip"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (the compiler is not allowed to form v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-pass-failed",
    "-I" + os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


HDF5_ROOT = os.environ.get("DNSBOX_HDF5_ROOT", "/opt/conda")
H5_LIB = os.path.join(LIBDIR, "libdnsbox_h5.so")


def build_h5(verbose: bool = False):
    """libdnsbox_h5.so (state files, host code) against the HDF5 C library, when one is present."""
    inc, libdir = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(libdir, "libhdf5.so"))):
        if verbose:
            print(f"no HDF5 C library under {HDF5_ROOT}: libdnsbox_h5.so not built", file=sys.stderr)
        return None
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "state_h5.cpp"), "-o", tmp, "-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir,
           # the C++ runtime linked in: the HDF5 directory's own libstdc++ must not be the one a
           # process loads (torch's libraries need a newer one)
           "-static-libstdc++", "-static-libgcc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for state_h5.cpp:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print("built", H5_LIB, file=sys.stderr)
    return H5_LIB


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, file=sys.stderr)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
