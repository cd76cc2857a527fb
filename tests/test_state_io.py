"""State files (dnsbox_amd.state_io over libdnsbox_h5.so): the reference's stateNNNN.h5 layout
(io.f90:100-278, 455-626), checked with the HDF5 command-line tools where they are installed."""
import os
import shutil
import subprocess

import numpy as np
import pytest

state_io = pytest.importorskip("dnsbox_amd.state_io")


def _lib_or_skip():
    if not os.path.exists(state_io.H5_LIB_PATH):
        pytest.skip("libdnsbox_h5.so was not built (no HDF5 C library on this host)")


def test_round_trip(tmp_path):
    _lib_or_skip()
    rng = np.random.default_rng(0)
    nz, ny, nx = 8, 6, 10
    u, v, w = (rng.standard_normal((nz, ny, nx)) for _ in range(3))
    path = tmp_path / "state0003.h5"
    state_io.write_state(path, u, v, w)
    assert state_io.state_dims(path) == (nx, ny, nz)
    got = state_io.read_state(path)
    assert all(np.array_equal(a, b) for a, b in zip(got, (u, v, w)))


def test_layout_is_fortran_nx_ny_nz(tmp_path):
    """u(i, j, k) of the Fortran array u(Nx, Ny, Nz) is element (k, j, i) of the C-order dataset
    {Nz, Ny, Nx}: h5dump shows the dataspace ( Nz, Ny, Nx ), IEEE FP64, x fastest in the data."""
    _lib_or_skip()
    h5dump = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
    if not h5dump:
        pytest.skip("h5dump not installed")
    nz, ny, nx = 3, 2, 4
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = (x + 10 * y + 100 * z).astype(np.float64)   # u(i, j, k) = i + 10 j + 100 k (0-based)
    path = tmp_path / "state0000.h5"
    state_io.write_state(path, u, 2 * u, 3 * u)
    out = subprocess.run([h5dump, "-d", "/u", str(path)], capture_output=True, text=True, check=True).stdout
    assert "H5T_IEEE_F64LE" in out
    assert f"SIMPLE {{ ( {nz}, {ny}, {nx} ) / ( {nz}, {ny}, {nx} ) }}" in out
    data = out[out.index("DATA {"):]
    # the first row printed is (0,0,:): x = 0..3 at y = z = 0
    assert "(0,0,0): 0, 1, 2, 3," in data
    names = subprocess.run([h5dump, "-n", str(path)], capture_output=True, text=True, check=True).stdout
    for ds in ("/u", "/v", "/w"):
        assert f"dataset    {ds}" in names


def test_wrong_shape_and_missing_file_are_errors(tmp_path):
    _lib_or_skip()
    path = tmp_path / "s.h5"
    state_io.write_state(path, *(np.zeros((4, 4, 4)) for _ in range(3)))
    import ctypes as C
    lib = state_io.load()
    buf = [np.zeros(8 * 8 * 8) for _ in range(3)]
    rc = lib.dnsbox_h5_read_state(os.fsencode(str(path)), 8, 8, 8,
                                  *[b.ctypes.data_as(C.POINTER(C.c_double)) for b in buf])
    assert rc == -1 and b"expected (8, 8, 8)" in lib.dnsbox_h5_last_error()
    with pytest.raises(Exception, match="cannot open"):
        state_io.read_state(tmp_path / "missing.h5")
