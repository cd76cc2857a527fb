"""Pins the CPU oracle (oracle/dnsbox_oracle.c) before it is trusted as the checker.

The reference holds no golden vectors or fixtures for this path (its only
checks are the Shapiro known-answer run, state.f90:287-348, and the FFT round
trip of single/test.f90:12-31).  Each oracle function is pinned here by
library results (numpy.fft), closed forms, invariants, or an independent
formulation, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails one of them.
"""
import numpy as np
import pytest

from oracle_lib import Oracle, to_torch_layout

ALL_N = [16, 32, 48]


def spec_k(o):
    """table wavenumbers broadcast to the oracle spectral shape (nxh, ny, nz)."""
    kx = o.ktable(o.nx)[: o.nx // 2 + 1]
    ky = o.ktable(o.ny)
    kz = o.ktable(o.nz)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def hermitian_weights(o):
    w = np.full(o.spec_shape, 2.0)
    w[0] = 1.0
    w[o.nx // 2] = 1.0
    return w


# --------------------------------------------------------------------------- k-tables
def test_ktable_reference_values():
    # variables.f90:226-262: 0..N/2-1, Nyquist -> 0, then -(N/2-1)..-1
    o = Oracle(8)
    assert o.ktable(8).tolist() == [0, 1, 2, 3, 0, -3, -2, -1]
    assert o.ktable(6).tolist() == [0, 1, 2, 0, -2, -1]


@pytest.mark.parametrize("n,nspec", [(32, 17), (48, 26), (64, 35), (128, 72), (256, 146), (512, 294)])
def test_nspec(n, nspec):
    # variables.f90:117-122 (cubic truncation)
    assert Oracle(n).nspec() == nspec


def test_mask_rule():
    o = Oracle(32)
    # |k| >= 32/3 = 10.67 masked on any axis; |k| = 10 kept
    assert o.masked(11, 0, 1)
    assert not o.masked(10, 0, 1)
    assert o.masked(0, 32 - 11, 1)
    assert o.masked(0, 0, 0)  # k = 0
    # Nyquist index has table k = 0: kept in parity mode, removed otherwise
    assert not o.masked(16, 1, 0)
    assert Oracle(32, parity_mode=False).masked(16, 1, 0)
    # N = 48: threshold (24)*(2/3) rounds to exactly 16.0, so |k| = 16 is masked (".ge.")
    o48 = Oracle(48)
    assert o48.masked(16, 0, 1) and not o48.masked(15, 0, 1)


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 12, 12), (16, 16, 16), (48, 48, 48), (6, 8, 12)])
def test_r2c_matches_numpy(shape):
    nx, ny, nz = shape
    o = Oracle(nx, ny=ny, nz=nz)
    x = np.random.default_rng(1).standard_normal(o.phys_shape)
    ref = np.fft.rfftn(x)
    got = to_torch_layout(o.r2c(x))
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_r2c_brute_force_dft():
    nx, ny, nz = 6, 4, 5
    o = Oracle(nx, ny=ny, nz=nz)
    x = np.random.default_rng(2).standard_normal(o.phys_shape)  # [z][y][x]
    got = o.r2c(x)
    zz, yy, xx = np.meshgrid(range(nz), range(ny), range(nx), indexing="ij")
    for ix in range(nx // 2 + 1):
        for iy in range(ny):
            for iz in range(nz):
                ph = np.exp(-2j * np.pi * (ix * xx / nx + iy * yy / ny + iz * zz / nz))
                assert abs(got[ix, iy, iz] - np.sum(x * ph)) < 1e-12


@pytest.mark.parametrize("n", [8, 12, 32])
def test_roundtrip_and_parseval(n):
    # single/test.f90:12-31: backward(forward(u)) = N^3 u
    o = Oracle(n)
    x = np.random.default_rng(3).standard_normal(o.phys_shape)
    s = o.r2c(x)
    assert np.abs(o.c2r(s) / n ** 3 - x).max() < 1e-13
    w = hermitian_weights(o)
    assert abs(np.sum(w * np.abs(s) ** 2) / n ** 3 - np.sum(x ** 2)) < 1e-10 * np.sum(x ** 2)


def test_c2r_matches_numpy_irfftn():
    o = Oracle(16)
    x = np.random.default_rng(4).standard_normal(o.phys_shape)
    s = o.r2c(x)
    ref = np.fft.irfftn(to_torch_layout(s), s=o.phys_shape, axes=(0, 1, 2)) * 16 ** 3
    assert np.abs(o.c2r(s) - ref).max() < 1e-11


def test_fft1d_radix3_and_prime():
    for n, sign in [(24, -1), (48, 1), (7, -1), (30, -1)]:
        x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
        y = x.copy()
        import ctypes as C
        o = Oracle(8)
        o.L.orc_fft1d(y.ctypes.data_as(C.POINTER(C.c_double)), n, sign)
        ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
        assert np.abs(y - ref).max() < 1e-12 * max(1, np.abs(ref).max())


# --------------------------------------------------------------------------- fields
def random_field(o, seed=0, k0=4.0):
    return o.init_solenoidal(seed, 0, k0, 0.5)


@pytest.mark.parametrize("n", ALL_N)
def test_init_solenoidal_properties(n):
    o = Oracle(n)
    f = random_field(o, 5)
    kx, ky, kz = spec_k(o)
    div = kx * f[0] + ky * f[1] + kz * f[2]
    assert np.abs(div).max() < 1e-13 * max(np.abs(a).max() for a in f)
    d, _ = o.diagnostics(f, 0.01)
    assert abs(d[0] - 0.5) < 1e-12
    # Hermitian on the kx = 0 plane: the field survives c2r -> r2c unchanged
    back = [o.r2c(o.c2r(a) / n ** 3) for a in f]
    assert max(np.abs(a - b).max() for a, b in zip(f, back)) < 1e-13 * max(np.abs(a).max() for a in f)
    # masked modes and Nyquist planes are empty
    m = o.mask_array()
    for a in f:
        assert np.all(a[m] == 0)
        assert np.all(a[n // 2] == 0) and np.all(a[:, n // 2] == 0) and np.all(a[:, :, n // 2] == 0)
    # deterministic in the seed, different across seeds and boxes
    assert all(np.array_equal(a, b) for a, b in zip(f, random_field(o, 5)))
    assert not np.allclose(f[0], random_field(o, 6)[0])
    assert not np.allclose(f[0], o.init_solenoidal(5, 1, 4.0, 0.5)[0])


def test_init_spectrum_shape():
    # shell energy follows k^4 exp(-2 (k/k0)^2) up to shell-counting noise
    o = Oracle(64)
    f = o.init_solenoidal(1, 0, 4.0, 0.5)
    _, es = o.diagnostics(f, 0.01)
    k = np.arange(1, len(es) + 1)
    model = k ** 4.0 * np.exp(-2 * (k / 4.0) ** 2)
    sel = slice(2, 10)
    ratio = es[sel] / model[sel]
    assert ratio.std() / ratio.mean() < 0.25


# --------------------------------------------------------------------------- RHS
def numpy_convective_rhs(o, f):
    """-P(u . grad u) computed with numpy.fft in the convective form (independent of the
    oracle's rotational-form code path)."""
    n = o.nx
    kx, ky, kz = spec_k(o)
    ks = [kx, ky, kz]
    tl = [to_torch_layout(a) for a in f]
    kt = [to_torch_layout(k) for k in ks]
    u = [np.fft.irfftn(a, s=o.phys_shape, axes=(0, 1, 2)) for a in tl]
    grads = [[np.fft.irfftn(1j * kt[j] * tl[i], s=o.phys_shape, axes=(0, 1, 2)) for j in range(3)] for i in range(3)]
    conv = [sum(u[j] * grads[i][j] for j in range(3)) for i in range(3)]
    ch = [np.fft.rfftn(c) for c in conv]
    k2 = sum(k * k for k in kt)
    k2[k2 == 0] = 1
    kc = sum(kt[j] * ch[j] for j in range(3)) / k2
    return [np.transpose(-(ch[i] - kt[i] * kc), (2, 1, 0)) for i in range(3)]


@pytest.mark.parametrize("n", [16, 32])
def test_rhs_rotational_equals_convective(n):
    """For a 2/3-truncated solenoidal field, P(u x w) = -P(u . grad u) on every retained mode
    off the Nyquist planes (the two forms differ by a gradient, which the projection removes;
    the truncation keeps retained modes alias-free)."""
    o = Oracle(n, nu=0.05)
    f = random_field(o, 9)
    got = o.rhs(f)
    ref = numpy_convective_rhs(o, f)
    m = o.mask_array()
    keep = ~m
    keep[n // 2, :, :] = False
    keep[:, n // 2, :] = False
    keep[:, :, n // 2] = False
    scale = max(np.abs(r[keep]).max() for r in ref)
    for g, r in zip(got, ref):
        assert np.abs(g[keep] - r[keep]).max() < 1e-12 * scale


@pytest.mark.parametrize("n", ALL_N)
def test_rhs_invariants(n):
    o = Oracle(n, nu=0.05)
    f = random_field(o, 2)
    nl = o.rhs(f)
    kx, ky, kz = spec_k(o)
    scale = max(np.abs(a).max() for a in nl)
    # divergence free
    assert np.abs(kx * nl[0] + ky * nl[1] + kz * nl[2]).max() < 1e-12 * scale * n
    # masked modes are zero
    m = o.mask_array()
    assert all(np.all(a[m] == 0) for a in nl)
    # energy conservation of the (unforced) nonlinear term
    w = hermitian_weights(o)
    dedt = sum(np.sum(w * np.real(np.conj(a) * b)) for a, b in zip(f, nl))
    mag = sum(np.sum(w * np.abs(a) * np.abs(b)) for a, b in zip(f, nl))
    assert abs(dedt) < 1e-12 * mag


def test_rhs_shapiro_vanishes():
    o = Oracle(32, nu=0.05)
    u = o.init_shapiro(0.0)
    f = [o.r2c(a) for a in u]
    nl = o.rhs(f)
    assert max(np.abs(a).max() for a in nl) < 1e-12 * max(np.abs(a).max() for a in f)


def test_kolmogorov_forcing_closed_form():
    n = 32
    o = Oracle(n, kolmo_amp=1.0, kolmo_k=4)
    f = [np.zeros(o.spec_shape, complex) for _ in range(3)]
    nl = o.rhs(f)
    expect = np.zeros(o.spec_shape, complex)
    expect[0, 4, 0] = -0.5j * n ** 3
    expect[0, n - 4, 0] = 0.5j * n ** 3
    assert np.abs(nl[0] - expect).max() < 1e-9
    assert np.abs(nl[1]).max() == 0 and np.abs(nl[2]).max() == 0
    # physical space: f = sin(4 y) on the grid y_j = -pi + 2 pi j / N
    phys = o.c2r(nl[0]) / n ** 3
    y = -np.pi + 2 * np.pi * np.arange(n) / n
    assert np.abs(phys - np.sin(4 * y)[None, :, None]).max() < 1e-13


def test_band_forcing_power_input():
    """rhs.f90:109-118: the band term injects exactly Pin into Ekin."""
    n = 32
    Pin = 0.1
    ob = Oracle(n, nu=0.01, Pin=Pin, kcut=2.5, bandlim=True)
    o0 = Oracle(n, nu=0.01)
    f = random_field(ob, 4, k0=2.0)
    dn = [a - b for a, b in zip(ob.rhs(f), o0.rhs(f))]
    w = hermitian_weights(ob) / 2.0
    power = 2.0 * sum(np.sum(w * np.real(np.conj(a) * b)) for a, b in zip(f, dn)) / float(n) ** 6
    assert abs(power - Pin) < 1e-12
    # Eband is the state_kinetic energy restricted to |k| <= kcut (w = 1/2 on kx = 0)
    kx, ky, kz = spec_k(ob)
    band = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2) <= 2.5
    eb = sum(np.sum((w * np.abs(a) ** 2)[band]) for a in f) / float(n) ** 6
    assert abs(ob.eband(f) - eb) < 1e-14


# --------------------------------------------------------------------------- stepping
def test_int_fact_closed_form():
    o = Oracle(16, nu=0.03, Q=0.2)
    itf = o.int_fact(0.01)
    kx, ky, kz = spec_k(o)
    assert np.abs(itf - np.exp((-0.03 * (kx ** 2 + ky ** 2 + kz ** 2) + 0.2) * 0.01)).max() < 1e-15


def test_heun_shapiro_kat():
    # state.f90:309-348: error vs exp(-3 nu t) at the end of the run
    o = Oracle(16, nu=0.05)
    u = o.init_shapiro(0.0)
    f = o.project(o.dealias([o.r2c(a) for a in u]))
    g = o.heun(f, 100, 0.01)
    phys = [o.c2r(a) / 16 ** 3 for a in g]
    assert o.shapiro_error(1.0, phys) < 1e-13


def _err(a, b):
    return np.sqrt(sum(np.sum(np.abs(x - y) ** 2) for x, y in zip(a, b)))


@pytest.mark.parametrize("stepper", ["heun", "ab2"])
def test_second_order_convergence(stepper):
    o = Oracle(16, nu=0.05, kolmo_amp=1.0, kolmo_k=2)
    f = random_field(o, 3, k0=3.0)
    T = 0.08
    run = o.heun if stepper == "heun" else o.ab2
    fine = o.heun(f, 64, T / 64)
    errs = [_err(run(f, m, T / m), fine) for m in (4, 8, 16)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 4.8 and 3.3 < r2 < 4.8, (errs, r1, r2)


def test_ab2_first_step_is_heun():
    o = Oracle(16, nu=0.05)
    f = random_field(o, 8)
    a = o.ab2(f, 1, 0.01)
    h = o.heun(f, 1, 0.01)
    assert _err(a, h) == 0.0


def test_project_and_dealias_idempotent():
    o = Oracle(16)
    rng = np.random.default_rng(0)
    f = [o.r2c(rng.standard_normal(o.phys_shape)) for _ in range(3)]
    p1 = o.project(f)
    p2 = o.project(p1)
    assert _err(p1, p2) < 1e-12 * _err(p1, [0 * a for a in p1])
    kx, ky, kz = spec_k(o)
    assert np.abs(kx * p1[0] + ky * p1[1] + kz * p1[2]).max() < 1e-10
    d1 = o.dealias(f)
    assert _err(d1, o.dealias(d1)) == 0.0


# --------------------------------------------------------------------------- diagnostics
def test_diagnostics_closed_forms():
    n = 32
    nu = 0.01
    o = Oracle(n, nu=nu, Q=0.3)
    f = random_field(o, 12)
    d, es = o.diagnostics(f, 0.005)
    u = [o.c2r(a) / n ** 3 for a in f]
    # Ekin = 1/2 <|u|^2> (Parseval)
    assert abs(d[0] - 0.5 * np.mean(sum(a * a for a in u))) < 1e-14
    # Disp = nu <|grad u|^2> via numpy gradients of the physical field
    kx, ky, kz = spec_k(o)
    g2 = 0.0
    for a in f:
        for k in (kx, ky, kz):
            g2 += np.mean(np.fft.irfftn(to_torch_layout(1j * k * a), s=o.phys_shape, axes=(0, 1, 2)) ** 2)
    assert abs(d[1] - nu * g2) < 1e-12 * d[1]
    assert abs(d[2] - 2 * 0.3 * d[0]) < 1e-15
    # Courant = dt sum_i max|u_i| / dx
    c = 0.005 * sum(np.abs(a).max() for a in u) / (2 * np.pi / n)
    assert abs(d[3] - c) < 1e-12
    # divergence of a solenoidal field is rounding-level
    assert abs(d[4]) < 1e-12


def test_spectrum_single_mode():
    """state_spectrum: energy of the mode (kx=3, ky=4, kz=0) lands in shell 5 with weight 1/N^6
    and no kx-plane factor; a kx = 0 mode also has weight 1 (the double count vs Ekin)."""
    n = 32
    o = Oracle(n)
    f = [np.zeros(o.spec_shape, complex) for _ in range(3)]
    f[2][3, 4, 0] = 2.0 + 1.0j
    d, es = o.diagnostics(f, 0.01)
    assert abs(es[4] - 5.0 / n ** 6) < 1e-25 and np.count_nonzero(es) == 1
    assert abs(d[0] - 5.0 / n ** 6) < 1e-25
    g = [np.zeros(o.spec_shape, complex) for _ in range(3)]
    g[0][0, 1, 1] = 1.0
    d2, es2 = o.diagnostics(g, 0.01)
    assert abs(es2[0] - 1.0 / n ** 6) < 1e-25  # |k| = sqrt(2) rounds to shell 1
    assert abs(d2[0] - 0.5 / n ** 6) < 1e-25   # Ekin halves the kx = 0 plane
    assert d2[7] == 1.0                         # Esqrt2 (absk - sqrt2 < 1e-9 as written)
    assert d2[6] == 0.0                         # EOne


def test_cfl_controller():
    # dnsboxheun.f90:227-246 set_time_step, triggered as at :153-157
    o = Oracle(8)
    assert o.cfl_dt(0.3, 0.01, 0.15, 0.2, 1.0) == pytest.approx(0.01 * 0.175 / 0.3)
    assert o.cfl_dt(0.17, 0.01, 0.15, 0.2, 1.0) == 0.01        # inside the band: unchanged
    assert o.cfl_dt(0.1, 0.01, 0.15, 0.2, 0.012) == 0.012      # capped at tStepMax
    assert o.cfl_dt(0.1, 0.02, 0.15, 0.2, 0.02) == 0.02        # already at tStepMax
