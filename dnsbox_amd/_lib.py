"""ctypes binding of libdnsbox_hip.so (include/dnsbox.h).

The library is the product: when it is missing or fails to load, every entry
point raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libdnsbox_hip.so")

(D_EKIN, D_DISP, D_PROD, D_COURANT, D_DIVMAX, D_EZERO, D_EONE, D_ESQRT2, D_EBAND, D_TIME,
 D_DT, NDIAG) = range(12)
PASS_NAMES = ("zinv_curl", "yinv6", "xcross", "yfwd3", "zpred_zinv", "zcorr_zinv")
# launches of each pass per IF-Heun step once the step pipeline is running (the z-inverse of
# the first RHS is fused into the previous step's last pass)
PASS_PER_STEP = (0, 2, 2, 2, 1, 1)
DIAG_NAMES = ("Ekin", "Disp", "Prod", "Courant", "divMax", "EZero", "EOne", "Esqrt2", "Eband",
              "time", "dt")


class Params(C.Structure):
    """dnsbox_params (include/dnsbox.h)."""
    _fields_ = [
        ("nu", C.c_double), ("Q", C.c_double), ("Pin", C.c_double), ("kcut", C.c_double),
        ("kolmo_amp", C.c_double), ("deltak", C.c_double), ("cfl_min", C.c_double),
        ("cfl_max", C.c_double), ("dt_max", C.c_double), ("dt", C.c_double),
        ("n", C.c_int), ("precision", C.c_int), ("kolmo_k", C.c_int), ("bandlim", C.c_int),
        ("spherical", C.c_int), ("parity_mode", C.c_int), ("adaptive_dt", C.c_int),
        ("state_precision", C.c_int), ("alpha_x", C.c_double), ("alpha_y", C.c_double),
        ("alpha_z", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
    ]


class DnsboxError(RuntimeError):
    pass


# exported symbols and their signatures: (restype, argtypes)
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "dnsbox_last_error": (C.c_char_p, []),
    "dnsbox_version": (C.c_char_p, []),
    "dnsbox_params_default": (C.c_int, [C.POINTER(Params)]),
    "dnsbox_create": (C.c_int, [C.POINTER(Params), C.POINTER(_vp)]),
    "dnsbox_destroy": (C.c_int, [_vp]),
    "dnsbox_set_stream": (C.c_int, [_vp, _vp]),
    "dnsbox_sync": (C.c_int, [_vp]),
    "dnsbox_nspec": (C.c_int, [_vp]),
    "dnsbox_set_state": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dnsbox_get_state": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dnsbox_set_state_physical": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dnsbox_get_state_physical": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dnsbox_init_random": (C.c_int, [_vp, C.c_uint64, C.c_int, C.c_double, C.c_double]),
    "dnsbox_project": (C.c_int, [_vp, C.c_int, C.c_int]),
    "dnsbox_int_fact": (C.c_int, [_vp, C.c_double]),
    "dnsbox_rhs_nonlinear": (C.c_int, [_vp]),
    "dnsbox_get_nonlin": (C.c_int, [_vp, _dp, _dp, _dp]),
    "dnsbox_step": (C.c_int, [_vp, C.c_int, C.c_int]),
    "dnsbox_diagnostics": (C.c_int, [_vp, _dp, _dp]),
    "dnsbox_get_time": (C.c_int, [_vp, _dp, _dp]),
    "dnsbox_set_time": (C.c_int, [_vp, C.c_double, C.c_double]),
    "dnsbox_time_passes": (C.c_int, [_vp, C.c_int, _dp]),
    "dnsbox_state_ptr": (C.c_int, [_vp, C.POINTER(_vp)]),
    "dnsbox_pass_bytes": (C.c_int, [_vp, _dp]),
    "dnsbox_step_dev_record": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    "dnsbox_step_vjp_pred_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double]),
    "dnsbox_step_dev_record_mixed": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp]),
    "dnsbox_step_vjp_mixed_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp, C.c_int]),
    "dnsbox_step_vjp_mixed_seq_dev": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _vp, _vp, C.c_double,
                                                C.POINTER(_vp), C.c_int, C.c_int]),
    "dnsbox_set_pruning": (C.c_int, [_vp, C.c_int]),
    "dnsbox_get_pruning": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "dnsbox_set_vjp_cube": (C.c_int, [_vp, C.c_int]),
    "dnsbox_get_last_step_pruned": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "dnsbox_group_set_vjp_cube": (C.c_int, [_vp, C.c_int]),
    "dnsbox_group_get_last_step_pruned": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "dnsbox_rhs_dev": (C.c_int, [_vp, _vp, _vp]),
    "dnsbox_step_dev": (C.c_int, [_vp, _vp, C.c_int]),
    "dnsbox_rhs_vjp_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "dnsbox_step_vjp_dev": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double]),
    "dnsbox_rhs_adj_dev": (C.c_int, [_vp, _vp, _vp, C.c_int]),
    "dnsbox_adj_descent_step_dev": (C.c_int, [_vp, _vp, C.c_double, C.c_int, C.c_int]),
    "dnsbox_adj_descent_step": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int]),
    "dnsbox_rhs_dev_noalloc": (C.c_int, [C.POINTER(Params), _vp, _vp, _vp]),
    "dnsbox_group_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(_vp)]),
    "dnsbox_group_destroy": (C.c_int, [_vp]),
    "dnsbox_group_size": (C.c_int, [_vp]),
    "dnsbox_group_box": (C.c_int, [_vp, C.c_int, C.POINTER(_vp)]),
    "dnsbox_group_step": (C.c_int, [_vp, C.c_int]),
    "dnsbox_group_sync": (C.c_int, [_vp]),
    "dnsbox_group_set_stream": (C.c_int, [_vp, _vp]),
    "dnsbox_group_step_dev": (C.c_int, [_vp, C.POINTER(_vp), C.c_int]),
    "dnsbox_group_step_dev_record": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(_vp), C.POINTER(_vp)]),
    "dnsbox_group_step_vjp_pred_dev": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                                C.POINTER(_vp), C.POINTER(C.c_double)]),
    "dnsbox_group_step_dev_record_mixed": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(_vp), C.POINTER(_vp),
                                                    C.POINTER(_vp)]),
    "dnsbox_group_step_vjp_mixed_dev": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                                 C.POINTER(_vp), C.POINTER(C.c_double), C.POINTER(_vp), C.c_int]),
    "dnsbox_group_step_vjp_dev": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), _dp]),
}

_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64.so.7.  Importing torch
    # first makes the library bind to that already-loaded runtime (same soname), so torch
    # tensors and the library share one runtime whichever is used first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise DnsboxError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        msg = load().dnsbox_last_error().decode(errors="replace")
        raise DnsboxError(f"dnsbox error {rc}: {msg}")
    return rc
