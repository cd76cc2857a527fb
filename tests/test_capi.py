"""The C-ABI library loads and exports every entry point include/dnsbox.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# header -> the library that exports its functions
HEADERS = {"dnsbox.h": "libdnsbox_hip.so", "dnsbox_io.h": "libdnsbox_h5.so"}


def declared_functions(header="dnsbox.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dnsbox_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_every_header_maps_to_a_library():
    assert sorted(h for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h")) == sorted(HEADERS)


def test_state_file_library_exports_its_header():
    path = os.path.join(ROOT, "dnsbox_amd", "lib", HEADERS["dnsbox_io.h"])
    if not os.path.exists(path):
        pytest.skip("libdnsbox_h5.so was not built (no HDF5 C library on this host)")
    from dnsbox_amd import state_io
    lib = state_io.load()
    names = declared_functions("dnsbox_io.h")
    assert len(names) == 4 and not [n for n in names if not hasattr(lib, n)]


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
    # 10 doubles, 8 ints, 3 doubles (alpha_x/y/z), 3 ints (nx, ny, nz), padded to 8 bytes
    assert C.sizeof(Params) == 10 * 8 + 8 * 4 + 3 * 8 + 3 * 4 + 4
    assert Params.alpha_x.offset == 112 and Params.nx.offset == 136


def test_grid_parameters_checked_before_device_work():
    """The grid (parameters.f90:12-14) and box lengths (:18-20) are checked before any device work:
    the cubic grids and the listed non-cubic grids pass (here, without a GPU, the call then fails
    with DNSBOX_E_HIP), any other grid, a negative length or nx, a Kolmogorov wavenumber at or past
    ny/2 and a non-cubic group member fail with DNSBOX_E_ARG."""
    from dnsbox_amd import _lib
    lib = _lib.load()
    p = _lib.Params()
    assert lib.dnsbox_params_default(C.byref(p)) == 0
    assert (p.alpha_x, p.alpha_y, p.alpha_z, p.nx, p.ny, p.nz) == (1.0, 1.0, 1.0, 0, 0, 0)
    h = C.c_void_p()
    for n, nx, ny, nz in [(32, 0, 0, 64), (32, 0, 48, 16), (48, 0, 16, 32), (64, 0, 0, 128), (256, 0, 0, 128),
                          (0, 32, 32, 64), (128, 0, 0, 0)]:
        p.n, p.nx, p.ny, p.nz = n, nx, ny, nz
        assert lib.dnsbox_create(C.byref(p), C.byref(h)) == -2, (n, nx, ny, nz)   # DNSBOX_E_HIP: no device
        assert b"no HIP device" in lib.dnsbox_last_error()
    for n, nx, ny, nz in [(32, 0, 0, 128), (64, 0, 0, 32), (32, 48, 32, 0), (24, 0, 0, 0), (0, 32, 32, 0)]:
        p.n, p.nx, p.ny, p.nz = n, nx, ny, nz
        assert lib.dnsbox_create(C.byref(p), C.byref(h)) == -1, (n, nx, ny, nz)
        msg = lib.dnsbox_last_error()
        assert b"grid" in msg and (b"32x48x16" in msg or b"cubic grids" in msg), msg
    p.n, p.nx, p.ny, p.nz = 32, -1, 0, 0
    assert lib.dnsbox_create(C.byref(p), C.byref(h)) == -1
    p.nx, p.ny, p.nz, p.kolmo_amp, p.kolmo_k = 0, 48, 16, 1.0, 24   # forcing sin(k y) needs k < ny/2
    assert lib.dnsbox_create(C.byref(p), C.byref(h)) == -1
    assert b"kolmo_k" in lib.dnsbox_last_error()
    p.kolmo_k = 23
    assert lib.dnsbox_create(C.byref(p), C.byref(h)) == -2
    p.kolmo_amp, p.kolmo_k = 0.0, 0
    p.ny, p.nz, p.alpha_y = 0, 0, -1.0
    assert lib.dnsbox_create(C.byref(p), C.byref(h)) == -1
    assert b"alpha" in lib.dnsbox_last_error()
    p.alpha_y, p.nz = 1.0, 64
    arr = (_lib.Params * 1)(p)
    g = C.c_void_p()
    assert lib.dnsbox_group_create(arr, 1, C.byref(g)) == -1
    assert b"cubic" in lib.dnsbox_last_error()


def test_step_map_refuses_adaptive_boxes():
    """The step adjoints differentiate fixed-dt steps; a box with the CFL controller on (the
    reference default, tStepFix = .false.) is refused before any step is taken."""
    import types

    torch = pytest.importorskip("torch")
    from dnsbox_amd.adjoint import group_step_map, step_map
    box = types.SimpleNamespace(params=types.SimpleNamespace(adaptive_dt=1))
    with pytest.raises(ValueError, match="adaptive_dt=0"):
        step_map(box, torch.zeros(3, 4, 4, 3, dtype=torch.complex128), 2, 0.01, 1)
    grp = types.SimpleNamespace(boxes=[types.SimpleNamespace(params=types.SimpleNamespace(adaptive_dt=0)), box])
    with pytest.raises(ValueError, match="adaptive_dt=0"):
        group_step_map(grp, [torch.zeros(3, 4, 4, 3, dtype=torch.complex128)] * 2, 2, [0.01, 0.01], 1)


def test_errors_without_device_do_not_abort():
    """Argument errors are reported through return codes and dnsbox_last_error."""
    from dnsbox_amd import _lib
    lib = _lib.load()
    p = _lib.Params()
    assert lib.dnsbox_params_default(C.byref(p)) == 0
    p.n = 20
    h = C.c_void_p()
    rc = lib.dnsbox_create(C.byref(p), C.byref(h))
    assert rc == -1 and b"unsupported grid 20 x 20 x 20" in lib.dnsbox_last_error()
    assert lib.dnsbox_step(None, 1, 1) == -1
    # the segment step adjoint refuses a missing context before it reads its pointer arrays
    arr = (C.c_void_p * 2)()
    assert lib.dnsbox_step_vjp_mixed_seq_dev(None, arr, arr, None, None, 0.01, arr, 1, 1) == -1
