"""The C-ABI library loads and exports every entry point include/dnsbox.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dnsbox_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("dnsbox_create", "dnsbox_rhs_nonlinear", "dnsbox_step", "dnsbox_int_fact",
                     "dnsbox_diagnostics", "dnsbox_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from dnsbox_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from dnsbox_amd import _lib
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_params_struct_layout():
    from dnsbox_amd._lib import Params
    # 10 doubles then 8 ints, natural alignment, no padding
    assert C.sizeof(Params) == 10 * 8 + 8 * 4


def test_errors_without_device_do_not_abort():
    """Argument errors are reported through return codes and dnsbox_last_error."""
    from dnsbox_amd import _lib
    lib = _lib.load()
    p = _lib.Params()
    assert lib.dnsbox_params_default(C.byref(p)) == 0
    p.n = 20
    h = C.c_void_p()
    rc = lib.dnsbox_create(C.byref(p), C.byref(h))
    assert rc == -1 and b"unsupported n" in lib.dnsbox_last_error()
    assert lib.dnsbox_step(None, 1, 1) == -1
