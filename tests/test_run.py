"""Run control around the step (dnsbox_amd/run.py): file formats of io.f90, the time loop of
dnsboxheun.f90, and the reference's python/analysis.py reading a run made here."""
import os

import numpy as np
import pytest

from oracle_lib import Oracle


def test_fortran_edit_descriptors():
    from dnsbox_amd.run import fortran_e, fortran_f
    assert fortran_e(1.0) == "  0.100000000000E+01"
    assert fortran_e(-0.0123456789012345) == " -0.123456789012E-01"
    assert fortran_e(0.0) == "  0.000000000000E+00"
    assert fortran_e(9.9999999999999e-5) == "  0.100000000000E-03"
    assert len(fortran_e(123456.0)) == 20
    assert fortran_f(0.05) == " 0.050000000"
    for x in (1.0, -3.25e-7, 6.02e23, 0.5):
        assert float(fortran_e(x)) == pytest.approx(x, rel=1e-11)


def test_info_file_reads_like_the_reference(tmp_path):
    """python/analysis.py:12-21 reads alpha (line 4), nu (line 6) and Q (line 7) from info.dat."""
    from types import SimpleNamespace
    from dnsbox_amd.run import write_info
    p = SimpleNamespace(nu=0.02, Q=0.1, Pin=0.0, kcut=0.0, deltak=1.0, bandlim=0)
    box = SimpleNamespace(params=p, n=32, nx=32, ny=48, nz=16)
    path = tmp_path / "info.dat"
    write_info(str(path), box, 1000, 10, 0.005, 0.0)
    lines = path.read_text().splitlines()
    assert lines[0] == "Nx,y,z =   32  48  16"
    assert [float(x) for x in lines[3].split()[2:5]] == [1.0, 1.0, 1.0]
    assert float(lines[5].split()[2]) == 0.02 and float(lines[6].split()[2]) == 0.1


@pytest.mark.gpu
def test_driver_files_match_the_oracle(gpu_lib, tmp_path):
    """32^3 fixed-dt run of 40 steps, statistics every 10, spectrum/state every 20: Ekin.dat
    rows and spectra against the oracle stepped to the same times; stops at nt."""
    from dnsbox_amd import Box
    from dnsbox_amd.run import run
    n, dt, nu = 32, 0.005, 0.05
    o = Oracle(n, nu=nu, kolmo_amp=1.0, kolmo_k=4)
    f0 = o.init_solenoidal(4, 0, 4.0, 0.5)
    with Box(n=n, precision=64, nu=nu, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.set_state(*f0)
        b.rhs_int_fact(dt)
        steps, why = run(b, str(tmp_path), 40, isave1=20, isave2=10, tstep_fix=True, courant_max=10.0)
    assert steps == 40 and "maximum" in why and not os.path.exists(tmp_path / "RUNNING")
    ekin = np.loadtxt(tmp_path / "Ekin.dat")
    assert ekin.shape == (5, 8)
    f = f0
    for j, (row, nsteps) in enumerate(zip(ekin, (0, 10, 10, 10, 10))):
        f = o.heun(f, nsteps, dt) if nsteps else f
        od, _ = o.diagnostics(f, dt)
        assert abs(row[0] - 10 * j * dt) < 1e-12
        for col, i in ((1, 0), (2, 1), (3, 2), (4, 3)):
            assert abs(row[col] - od[i]) <= 1e-10 * max(abs(od[i]), 1e-12), (col, row[col], od[i])
    # saves at n = 0, 20, 40 and, as in the reference, once more when the run ends at n = 40
    specs = sorted(p.name for p in tmp_path.glob("spec*.dat"))
    assert specs == ["spec0000.dat", "spec0001.dat", "spec0002.dat", "spec0003.dat"]
    assert len(list(tmp_path.glob("state*.h5"))) == 4
    s2 = np.loadtxt(tmp_path / "spec0002.dat")
    f40 = o.heun(f0, 40, dt)
    _, osp = o.diagnostics(f40, dt)
    assert np.allclose(s2[:, 1], osp, rtol=1e-10, atol=1e-14)
    assert np.array_equal(np.loadtxt(tmp_path / "spec0003.dat"), s2)
    # the state file holds the physical velocity of the oracle's state at t = 0.2
    from dnsbox_amd.state_io import read_state
    st = read_state(tmp_path / "state0002.h5")
    ref = [o.c2r(a) / n ** 3 for a in f40]
    assert st[0].shape == (n, n, n)
    assert max(np.abs(a - b).max() for a, b in zip(st, ref)) < 1e-10 * max(np.abs(b).max() for b in ref)
    # a run started from it (io_load_state, dnsboxheun.f90:48) continues the same trajectory
    with Box(n=n, precision=64, nu=nu, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.rhs_int_fact(dt)
        run(b, str(tmp_path / "resume"), 10, isave1=10, isave2=10, tstep_fix=True, courant_max=10.0,
            initial_state=str(tmp_path / "state0002.h5"))
        got = b.get_state()
    want = o.heun(o.project(o.dealias([o.r2c(a) for a in st])), 10, dt)
    err = np.sqrt(sum(np.sum(np.abs(g - w_) ** 2) for g, w_ in zip(got, want)) / sum(np.sum(np.abs(w_) ** 2) for w_ in want))
    assert err < 1e-10, err
    assert (tmp_path / "info.dat").read_text().splitlines()[5].split()[2] == "0.050000000"


@pytest.mark.gpu
def test_driver_reference_dt_policy_and_courant_abort(gpu_lib, tmp_path):
    from dnsbox_amd import Box
    from dnsbox_amd.run import run
    with Box(n=32, precision=64, nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.002) as b:
        b.init_random(1, 0, 4.0, 0.5)
        b.rhs_int_fact(0.002)
        steps, why = run(b, str(tmp_path / "a"), 60, isave1=30, isave2=10, tstep_fix=False,
                         courant_min=0.15, courant_max=0.2, tstep_max=0.01)
        ek = np.loadtxt(tmp_path / "a" / "Ekin.dat")
        _, dt_end = b.time()
        # set_time_step after the initial statistics aims at Courant (0.15 + 0.2) / 2
        assert steps == 60 and np.all(np.abs(ek[1:, 4] - 0.175) < 0.03), ek[:, 4]
        assert dt_end != 0.002
    with Box(n=32, precision=64, nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.01) as b:
        b.init_random(1, 0, 4.0, 0.5)
        b.rhs_int_fact(0.01)
        steps, why = run(b, str(tmp_path / "b"), 60, isave1=30, isave2=10, tstep_fix=True, courant_max=0.2)
    assert steps == 1 and "Courant" in why   # the initial Courant number ends the run after one step
