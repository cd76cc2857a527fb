"""State files of the reference, stateNNNN.h5 (io_save_state, io.f90:100-278; io_load_state,
io.f90:455-626), through libdnsbox_h5.so (include/dnsbox_io.h, the HDF5 C library).

Arrays are the physical velocity u, v, w in FP64 with numpy shape (Nz, Ny, Nx), which is the
reference's Fortran u(Nx, Ny, Nz) and the layout of Box.get_state_physical.  When the library
was not built (no HDF5 C library on the build host) every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import DnsboxError

PKG = os.path.dirname(os.path.abspath(__file__))
H5_LIB_PATH = os.path.join(PKG, "lib", "libdnsbox_h5.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    try:   # torch's HIP and system libraries first, as dnsbox_amd._lib.load does
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(H5_LIB_PATH):
        raise DnsboxError(f"{H5_LIB_PATH} is missing: it is built against the HDF5 C library "
                          "(hdf5.h and libhdf5) when the build host has one; state files need it")
    lib = C.CDLL(H5_LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    lib.dnsbox_h5_last_error.restype = C.c_char_p
    lib.dnsbox_h5_last_error.argtypes = []
    lib.dnsbox_h5_write_state.restype = C.c_int
    lib.dnsbox_h5_write_state.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp]
    lib.dnsbox_h5_read_state.restype = C.c_int
    lib.dnsbox_h5_read_state.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp]
    lib.dnsbox_h5_state_dims.restype = C.c_int
    lib.dnsbox_h5_state_dims.argtypes = [C.c_char_p, ip, ip, ip]
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise DnsboxError(f"state file error {rc}: {load().dnsbox_h5_last_error().decode(errors='replace')}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def write_state(path, u, v, w):
    """Write datasets u, v, w (numpy (Nz, Ny, Nx) float64 == Fortran (Nx, Ny, Nz)) to path."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (u, v, w)]
    if len({a.shape for a in arrs}) != 1 or arrs[0].ndim != 3:
        raise ValueError("u, v, w must be three arrays of one 3-D shape (Nz, Ny, Nx)")
    nz, ny, nx = arrs[0].shape
    _check(load().dnsbox_h5_write_state(os.fsencode(path), nx, ny, nz, *[_dp(a) for a in arrs]))


def state_dims(path):
    """Fortran shape (Nx, Ny, Nz) of the state file's datasets."""
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    _check(load().dnsbox_h5_state_dims(os.fsencode(path), C.byref(nx), C.byref(ny), C.byref(nz)))
    return nx.value, ny.value, nz.value


def read_state(path):
    """(u, v, w) of a state file, numpy (Nz, Ny, Nx) float64."""
    nx, ny, nz = state_dims(path)
    out = [np.empty((nz, ny, nx), np.float64) for _ in range(3)]
    _check(load().dnsbox_h5_read_state(os.fsencode(path), nx, ny, nz, *[_dp(a) for a in out]))
    return out
