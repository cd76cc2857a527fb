"""Non-cubic grids (Nx, Ny, Nz different, parameters.f90:12-14) against the oracle: the grids the
library lists (two small grids with three different lengths, the 2 pi x 2 pi x 4 pi box at
32 x 32 x 64 and 64 x 64 x 128 with alpha_z = 1/2, and 256 x 256 x 128), through the reference
layout of the C-ABI: IF-Heun steps pruned and unpruned, the adaptive step, IF-AB2, the RHS,
diagnostics, the physical-space round trip, the seeded generator and the adjoints.  FP64 at
1e-11 (steps) / 1e-12, FP32 at FP32 rounding (2e-5)."""
import numpy as np
import pytest

from oracle_lib import Oracle, from_torch_layout, to_torch_layout

pytestmark = pytest.mark.gpu

SMALL = [((32, 48, 16), (1.0, 1.0, 1.0)), ((48, 16, 32), (1.0, 1.0, 1.0)), ((32, 32, 64), (1.0, 1.0, 0.5))]
TOL = {64: 1e-11, 32: 2e-5}


def rel_l2(a, b):
    num = sum(np.linalg.norm((x - y).ravel()) ** 2 for x, y in zip(a, b))
    return float(np.sqrt(num / sum(np.linalg.norm(y.ravel()) ** 2 for y in b)))


def oracle(shape, alpha, **kw):
    nx, ny, nz = shape
    return Oracle(nx, ny=ny, nz=nz, alpha=alpha, **kw)


def box(shape, alpha, prec, **kw):
    from dnsbox_amd import Box
    nx, ny, nz = shape
    kw.setdefault("bandlim", 0)
    kw.setdefault("adaptive_dt", 0)
    return Box(n=nx, nx=nx, ny=ny, nz=nz, precision=prec, alpha_x=alpha[0], alpha_y=alpha[1], alpha_z=alpha[2], **kw)


def dev(f, prec):
    import torch
    t = torch.from_numpy(np.stack([to_torch_layout(a) for a in f])).cuda()
    return (t if prec == 64 else t.to(torch.complex64)).contiguous()


def host(t):
    import torch
    a = t.detach().cpu().to(torch.complex128).numpy()
    return [from_torch_layout(a[i]) for i in range(3)]


@pytest.mark.parametrize("pruned", [True, False])
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape,alpha", SMALL)
def test_grid_heun_matches_oracle(gpu_lib, shape, alpha, prec, pruned):
    dt, steps = 0.004, 4
    o = oracle(shape, alpha, nu=0.02, kolmo_amp=1.0, kolmo_k=3)
    f = o.init_solenoidal(11, 0, 3.0, 0.5)
    ref = o.heun(f, steps, dt)
    with box(shape, alpha, prec, nu=0.02, kolmo_amp=1.0, kolmo_k=3, dt=dt) as b:
        assert b.spec_shape == o.spec_shape and b.phys_shape == o.phys_shape
        b.set_pruning(pruned)
        b.set_state(*f)
        assert b.pruning() == pruned
        b.rhs_int_fact(dt)
        b.step(steps)
        got = b.get_state()
        t, _ = b.time()
    err = rel_l2(got, ref)
    print(f"grid heun {shape} alpha={alpha} prec={prec} pruned={pruned}: rel_l2={err:.3e}", flush=True)
    assert abs(t - steps * dt) < 1e-14
    assert err < TOL[prec]


@pytest.mark.parametrize("shape,alpha", SMALL)
def test_grid_rhs_and_diagnostics_match_oracle(gpu_lib, shape, alpha):
    """rhs_nonlinear with band forcing, then Ekin, dissipation, the Courant number (dx_d = L_d / N_d),
    div u and the shell spectrum of the state."""
    o = oracle(shape, alpha, nu=0.02, Pin=0.1, kcut=2.5, bandlim=True)
    f = o.init_solenoidal(12, 0, 3.0, 0.5)
    ref = o.rhs(f)
    dt = 0.003
    od, osp = o.diagnostics(f, dt)
    with box(shape, alpha, 64, nu=0.02, Pin=0.1, kcut=2.5, bandlim=1, dt=dt) as b:
        assert b.nspec() == o.nspec()
        b.set_state(*f)
        got = b.rhs_nonlinear()
        d, sp = b.diagnostics()
    err = rel_l2(got, ref)
    print(f"grid rhs {shape}: rel_l2={err:.3e}", flush=True)
    assert err < 1e-12
    # Ekin, Disp, Prod, Courant, EZero, EOne, Esqrt2, Eband (the oracle's order)
    for i in (0, 1, 2, 3, 5, 6, 7, 8):
        assert abs(d[i] - od[i]) <= 1e-11 * max(abs(od[i]), 1e-300) + 1e-14, (i, d[i], od[i])
    assert abs(d[4] - od[4]) < 1e-12   # divMax: rounding level for a solenoidal field
    np.testing.assert_allclose(sp, osp, rtol=1e-11, atol=1e-300)


@pytest.mark.parametrize("shape,alpha", SMALL[:2])
def test_grid_adaptive_heun_and_ab2_match_oracle(gpu_lib, shape, alpha):
    """The CFL controller (Courant from the x pass's maxima with dx_d = L_d / N_d) and IF-AB2."""
    dt0, cmin, cmax, dtmax = 0.01, 0.15, 0.2, 0.02
    o = oracle(shape, alpha, nu=0.02, kolmo_amp=1.0, kolmo_k=3)
    f = o.init_solenoidal(13, 0, 3.0, 1.0)
    ref, dt_ref, t_ref = o.heun_adaptive(f, 6, dt0, cmin, cmax, dtmax)
    with box(shape, alpha, 64, nu=0.02, kolmo_amp=1.0, kolmo_k=3, adaptive_dt=1, dt=dt0, cfl_min=cmin,
             cfl_max=cmax, dt_max=dtmax) as b:
        b.set_state(*f)
        b.step(6)
        got = b.get_state()
        t, dt = b.time()
    err = rel_l2(got, ref)
    print(f"grid adaptive {shape}: rel_l2={err:.3e} t={t:.6f} ({t_ref:.6f}) dt={dt:.6e} ({dt_ref:.6e})", flush=True)
    assert err < 1e-11 and abs(t - t_ref) < 1e-12 and abs(dt - dt_ref) < 1e-13
    dt = 0.004
    ref = o.ab2(f, 5, dt)
    with box(shape, alpha, 64, nu=0.02, kolmo_amp=1.0, kolmo_k=3, dt=dt) as b:
        b.set_state(*f)
        b.rhs_int_fact(dt)
        b.step(5, stepper=2)
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"grid ab2 {shape}: rel_l2={err:.3e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape,alpha", SMALL)
def test_grid_adjoints_match_oracle(gpu_lib, shape, alpha, prec):
    """RHS and step VJPs (the linearisation state in the kept cube or not) and rhsAdj."""
    dt = 0.004
    o = oracle(shape, alpha, nu=0.02, kolmo_amp=1.0, kolmo_k=3)
    u = o.init_solenoidal(14, 0, 3.0, 0.5)
    rng = np.random.default_rng(5)
    lam = [rng.standard_normal(o.spec_shape) + 1j * rng.standard_normal(o.spec_shape) for _ in range(3)]
    refs = {"rhs_vjp": o.rhs_vjp(u, lam), "step_vjp": o.heun_vjp(u, lam, dt), "rhs_adj": o.rhs_adj(u)}
    got = {}
    with box(shape, alpha, prec, nu=0.02, kolmo_amp=1.0, kolmo_k=3, dt=dt) as b:
        out = b.empty_state()
        b.rhs_vjp_dev(dev(u, prec), dev(lam, prec), out)
        got["rhs_vjp"] = host(out)
        b.set_vjp_cube(True)
        b.step_vjp_dev(dev(u, prec), dev(lam, prec), out, dt)
        b.set_vjp_cube(False)
        got["step_vjp"] = host(out)
        b.rhs_adj_dev(dev(u, prec), out, 1)
        got["rhs_adj"] = host(out)
    for k, ref in refs.items():
        err = rel_l2(got[k], ref)
        print(f"grid {k} {shape} prec={prec}: rel_l2={err:.3e}", flush=True)
        assert err < (1e-11 if prec == 64 else TOL[32]), k


@pytest.mark.parametrize("shape,alpha", SMALL)
def test_grid_generator_and_physical_round_trip(gpu_lib, shape, alpha):
    """init_random is the oracle's seeded generator on the grid; state_uhat2u / state_u2uhat
    against the oracle's transforms; a projection leaves the generated field unchanged."""
    o = oracle(shape, alpha, nu=0.02)
    ref = o.init_solenoidal(21, 3, 3.0, 0.7)
    m = o.nx * o.ny * o.nz
    with box(shape, alpha, 64, nu=0.02) as b:
        b.init_random(21, 3, 3.0, 0.7)
        got = b.get_state()
        assert rel_l2(got, ref) < 1e-13
        phys = b.get_state_physical()
        want = [o.c2r(a) / m for a in ref]
        assert phys[0].shape == o.phys_shape
        assert max(np.abs(p - w).max() for p, w in zip(phys, want)) < 1e-13 * max(np.abs(w).max() for w in want)
        b.set_state_physical(*want)
        assert rel_l2(b.get_state(), ref) < 1e-13
        b.state_project()
        assert rel_l2(b.get_state(), ref) < 1e-13


def test_grid_long_box_64x64x128(gpu_lib):
    """The 2 pi x 2 pi x 4 pi box at 64 x 64 x 128 (alpha_z = 1/2): IF-Heun in both precisions and
    the shell spectrum after the steps."""
    shape, alpha, dt = (64, 64, 128), (1.0, 1.0, 0.5), 0.004
    o = oracle(shape, alpha, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(31, 0, 3.0, 0.5)
    ref = o.heun(f, 3, dt)
    _, osp = o.diagnostics(ref, dt)
    for prec, sp in ((64, 0), (32, 64)):
        with box(shape, alpha, prec, nu=0.02, kolmo_amp=1.0, kolmo_k=4, dt=dt, state_precision=sp) as b:
            b.set_state(*f)
            b.rhs_int_fact(dt)
            b.step(3)
            got = b.get_state()
            sp_got = b.spectrum()
        err = rel_l2(got, ref)
        print(f"grid 64x64x128 prec={prec}: rel_l2={err:.3e}", flush=True)
        assert err < TOL[prec]
        assert np.allclose(sp_got, osp, rtol=1e-9 if prec == 64 else 1e-5, atol=1e-14)


def test_grid_256x256x128_fp32(gpu_lib):
    """256 x 256 x 128 (the register x pass runs cubic boxes only: here the LDS-staged x pass), FP32
    transforms with the FP64 state: two IF-Heun steps against the oracle."""
    shape, alpha, dt = (256, 256, 128), (1.0, 1.0, 1.0), 0.002
    o = oracle(shape, alpha, nu=0.005, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(41, 0, 4.0, 0.5)
    ref = o.heun(f, 2, dt)
    with box(shape, alpha, 32, nu=0.005, kolmo_amp=1.0, kolmo_k=4, dt=dt, state_precision=64) as b:
        b.set_state(*f)
        assert b.pruning()
        b.rhs_int_fact(dt)
        b.step(2)
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"grid 256x256x128 fp32: rel_l2={err:.3e}", flush=True)
    assert err < TOL[32]


@pytest.mark.parametrize("record", ["all", "checkpoint"])
def test_grid_tstep_gradient_and_descent(gpu_lib, record):
    """The T-step map on the 2 pi x 2 pi x 4 pi box at 32 x 32 x 64 (alpha_z = 1/2): the residual
    gradient against central differences along seeded directions (1e-6 relative; the mean square
    normalised by Nx Ny Nz), and three adjoint-descent iterations (projection with the scaled
    wavenumbers) lowering J monotonically."""
    import torch
    from dnsbox_amd.adjoint import adjoint_descent, residual
    shape, alpha, dt, T, k = (32, 32, 64), (1.0, 1.0, 0.5), 0.005, 20, 6
    o = oracle(shape, alpha, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    with box(shape, alpha, 64, nu=0.02, kolmo_amp=1.0, kolmo_k=4, dt=dt) as b:
        u0 = dev(o.init_solenoidal(11, 0, 3.0, 0.5), 64)
        x = u0.clone().requires_grad_(True)
        J = residual(b, x, T, dt, k, record)
        (g,) = torch.autograd.grad(J, x)
        worst = 0.0
        for d in range(4):
            v = dev(o.init_solenoidal(100 + d, 0, 3.0, 0.5), 64)
            eps = 1e-6
            with torch.no_grad():
                jp = float(residual(b, u0 + eps * v, T, dt, k))
                jm = float(residual(b, u0 - eps * v, T, dt, k))
            fd = (jp - jm) / (2 * eps)
            an = float((g.conj() * v).real.sum())
            worst = max(worst, abs(fd - an) / abs(fd))
        print(f"grid T-step gradient ({record}): worst rel {worst:.2e}", flush=True)
        assert worst < 1e-6
        # the residual is the mean square of the physical difference (Parseval on the grid)
        with torch.no_grad():
            uT = u0.clone()
            b.rhs_int_fact(dt)
            b.step_dev(uT, T)
        diff = [np.fft.irfftn(to_torch_layout(a), s=o.phys_shape, axes=(0, 1, 2))   # u = irfftn(u_hat)
                for a in host(uT - u0)]
        j = float(J.detach())
        assert abs(j - 0.5 * np.mean(sum(p ** 2 for p in diff))) < 1e-12 * j
        _, hist = adjoint_descent(b, u0, T, dt, k=k, iters=3, record=record)
    print(f"grid descent ({record}): {hist}", flush=True)
    assert all(b_ <= a_ for a_, b_ in zip(hist, hist[1:])) and hist[-1] < hist[0]


def test_grid_run_files_and_resume(gpu_lib, tmp_path):
    """The run driver on the 32 x 48 x 16 grid: Ekin.dat rows against the oracle, the state file
    holding the physical velocity of shape (Nz, Ny, Nx) (Fortran u(Nx, Ny, Nz)), and a run resumed
    from it continuing the same trajectory."""
    from dnsbox_amd.run import run
    from dnsbox_amd.state_io import read_state, state_dims
    shape, alpha, dt = (32, 48, 16), (1.0, 1.0, 1.0), 0.004
    o = oracle(shape, alpha, nu=0.02, kolmo_amp=1.0, kolmo_k=3)
    f0 = o.init_solenoidal(61, 0, 3.0, 0.5)
    with box(shape, alpha, 64, nu=0.02, kolmo_amp=1.0, kolmo_k=3, dt=dt) as b:
        b.set_state(*f0)
        b.rhs_int_fact(dt)
        steps, why = run(b, str(tmp_path), 20, isave1=10, isave2=10, tstep_fix=True, courant_max=10.0)
    assert steps == 20
    assert (tmp_path / "info.dat").read_text().splitlines()[0] == "Nx,y,z =   32  48  16"
    ekin = np.loadtxt(tmp_path / "Ekin.dat")
    f = f0
    for j, nsteps in enumerate((0, 10, 10)):
        f = o.heun(f, nsteps, dt) if nsteps else f
        od, _ = o.diagnostics(f, dt)
        assert abs(ekin[j][1] - od[0]) <= 1e-10 * od[0] and abs(ekin[j][2] - od[1]) <= 1e-10 * od[1]
    path = tmp_path / "state0001.h5"
    assert state_dims(path) == shape
    st = read_state(path)
    assert st[0].shape == o.phys_shape
    m = o.nx * o.ny * o.nz
    f10 = o.heun(f0, 10, dt)
    ref = [o.c2r(a) / m for a in f10]
    assert max(np.abs(a - r).max() for a, r in zip(st, ref)) < 1e-10 * max(np.abs(r).max() for r in ref)
    with box(shape, alpha, 64, nu=0.02, kolmo_amp=1.0, kolmo_k=3, dt=dt) as b:
        b.rhs_int_fact(dt)
        run(b, str(tmp_path / "resume"), 10, isave1=10, isave2=10, tstep_fix=True, courant_max=10.0,
            initial_state=str(path))
        got = b.get_state()
    want = o.heun(o.project(o.dealias([o.r2c(a) for a in st])), 10, dt)
    assert rel_l2(got, want) < 1e-10
