"""Pins the oracle's box lengths (alpha_x/y/z = 2 pi / L), spherical truncation and
parity_mode = 0 before the GPU path is compared with it.

Box lengths enter the reference through the k-tables (variables.f90:226-262, k = i alpha),
the 2/3-rule thresholds (N/2)(2 alpha/3) (rhs.f90:80-85, state.f90:191-196), Nspec
(variables.f90:112-123) and the Courant number's dx = L/N (io.f90:664-676).  The checks
below come from the mathematics (an exactly decaying shear wave, Parseval, solenoidality in
the scaled wavenumbers, an independent convective-form right-hand side, index invariance of
the 2/3 rule) and from brute-force counts of the spherical rules, chosen so that a missing
alpha on one side of a formula fails one of them.
"""
import numpy as np
import pytest

from oracle_lib import Oracle, to_torch_layout

ALPHA = (1.0, 2.0, 0.5)   # a 2 pi x pi x 4 pi box


def spec_k(o):
    """scaled wavenumbers broadcast to the oracle spectral shape (nxh, ny, nz)."""
    a = o.alpha
    kx = o.ktable(o.nx)[: o.nx // 2 + 1] * a[0]
    ky = o.ktable(o.ny) * a[1]
    kz = o.ktable(o.nz) * a[2]
    return np.meshgrid(kx, ky, kz, indexing="ij")


def hermitian_weights(o):
    w = np.full(o.spec_shape, 2.0)
    w[0] = 1.0
    w[o.nx // 2] = 1.0
    return w


def table_k(i, n):
    return i if i < n // 2 else (0 if i == n // 2 else i - n)


# --------------------------------------------------------------------------- box lengths
def test_nspec_with_box_lengths():
    # variables.f90:117-122 with thresholds (N/2)(2 alpha/3) = 16, 16, 8 at N = 48: sqrt(576) = 24
    assert Oracle(48, alpha=(1.0, 1.0, 0.5)).nspec() == 23
    # spherical: int((N/2)(2 alpha_x/3) / Deltak) - 1 = 8 - 1
    assert Oracle(48, alpha=(0.5, 1.0, 1.0), spherical=True).nspec() == 7


@pytest.mark.parametrize("n", [32, 48])
def test_cubic_rule_is_index_invariant(n):
    """|i| alpha >= (N/2)(2 alpha/3) <=> |i| >= N/3: the cubic rule removes the same indices for
    every box length (the threshold scales with alpha as the wavenumbers do)."""
    a = Oracle(n).mask_array()
    b = Oracle(n, alpha=ALPHA).mask_array()
    assert np.array_equal(a, b)


def test_spherical_rule_uses_scaled_wavenumbers():
    # rhs.f90:74-78: |k|^2 >= ((N/2) 0.5 alpha_x)^2 with k scaled; N = 32, alpha_x = 1 -> 64
    o = Oracle(32, spherical=True, alpha=(1.0, 2.0, 1.0))
    assert not o.masked(7, 1, 0)      # 49 + (2 * 1)^2 = 53
    assert o.masked(7, 2, 0)          # 49 + (2 * 2)^2 = 65
    assert not Oracle(32, spherical=True).masked(7, 2, 0)   # unscaled: 49 + 4 = 53


def test_shear_wave_decays_at_the_scaled_rate():
    """u = (sin(kf alpha_y y), 0, 0) has (u . grad) u = 0, so IF-Heun reproduces
    exp(-nu (kf alpha_y)^2 t) to rounding: pins alpha in the integrating factor and the
    projection (the rotational-form term is a pure gradient along y)."""
    n, kf, nu, dt, steps = 16, 3, 0.05, 0.01, 40
    o = Oracle(n, nu=nu, alpha=ALPHA)
    f = o.zeros_spec()
    f[0][0, kf, 0] = -0.5j * n ** 3
    f[0][0, n - kf, 0] = 0.5j * n ** 3
    g = o.heun(f, steps, dt)
    ratio = g[0][0, kf, 0] / f[0][0, kf, 0]
    expect = np.exp(-nu * (kf * ALPHA[1]) ** 2 * steps * dt)
    assert abs(ratio - expect) < 1e-13
    assert max(np.abs(a).max() for a in g[1:]) < 1e-12 * n ** 3


def numpy_convective_rhs(o, f):
    """-P(u . grad u) with numpy.fft and the scaled wavenumbers (independent of the oracle's
    rotational-form code)."""
    ks = spec_k(o)
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


def test_rhs_with_box_lengths_equals_convective_form():
    n = 16
    o = Oracle(n, nu=0.05, alpha=ALPHA)
    f = o.init_solenoidal(9, 0, 3.0, 0.5)
    got = o.rhs(f)
    ref = numpy_convective_rhs(o, f)
    keep = ~o.mask_array()
    keep[n // 2, :, :] = False
    keep[:, n // 2, :] = False
    keep[:, :, n // 2] = False
    scale = max(np.abs(r[keep]).max() for r in ref)
    for g, r in zip(got, ref):
        assert np.abs(g[keep] - r[keep]).max() < 1e-12 * scale


def test_fields_and_projection_are_solenoidal_in_scaled_wavenumbers():
    n = 16
    o = Oracle(n, alpha=ALPHA)
    kx, ky, kz = spec_k(o)
    f = o.init_solenoidal(3, 0, 3.0, 0.5)
    scale = max(np.abs(a).max() for a in f)
    assert np.abs(kx * f[0] + ky * f[1] + kz * f[2]).max() < 1e-13 * scale * n
    d, _ = o.diagnostics(f, 0.01)
    assert abs(d[0] - 0.5) < 1e-12
    rng = np.random.default_rng(0)
    g = o.project([o.r2c(rng.standard_normal(o.phys_shape)) for _ in range(3)])
    assert np.abs(kx * g[0] + ky * g[1] + kz * g[2]).max() < 1e-10
    # unscaled wavenumbers would leave a divergence
    k0 = np.meshgrid(o.ktable(n)[: n // 2 + 1], o.ktable(n), o.ktable(n), indexing="ij")
    assert np.abs(k0[0] * g[0] + k0[1] * g[1] + k0[2] * g[2]).max() > 1.0


def test_diagnostics_with_box_lengths():
    """Dissipation nu <|grad u|^2> = nu (kf alpha_y)^2 / 2 and Courant dt max|u| N alpha_x / (2 pi)
    of the unit shear wave u = sin(kf alpha_y y) e_x (io.f90:664-676, dx = L/N)."""
    n, kf, nu, dt = 16, 3, 0.02, 0.004
    o = Oracle(n, nu=nu, alpha=ALPHA)
    f = o.zeros_spec()
    f[0][0, kf, 0] = -0.5j * n ** 3
    f[0][0, n - kf, 0] = 0.5j * n ** 3
    d, _ = o.diagnostics(f, dt)
    assert abs(d[0] - 0.25) < 1e-14
    assert abs(d[1] - nu * (kf * ALPHA[1]) ** 2 / 2) < 1e-12
    assert abs(d[3] - dt * 1.0 * n * ALPHA[0] / (2 * np.pi)) < 1e-12


def test_alpha_is_used_at_single_precision():
    # parameters.f90:18-20 declare alpha real(kind=4): 0.1 acts as float32(0.1)
    o = Oracle(16, nu=0.05, alpha=(1.0, 0.1, 1.0))
    f = o.zeros_spec()
    f[0][0, 2, 0] = 1.0
    itf = o.int_fact(1.0)
    k32 = 2 * float(np.float32(0.1))
    assert itf[0, 2, 0] == np.exp(-0.05 * k32 * k32)


# --------------------------------------------------------------------------- spherical, parity
def brute_unmasked(n, r2_of, keep_k0=False):
    """count of stored modes (ix <= N/2, all iy, iz) with table k != 0, inside the cubic rule
    and with r2_of(k2) true, by enumeration"""
    cnt = 0
    thr = n / 3.0
    for ix in range(n // 2 + 1):
        kx = table_k(ix, n)
        for iy in range(n):
            ky = table_k(iy, n)
            for iz in range(n):
                kz = table_k(iz, n)
                k2 = kx * kx + ky * ky + kz * kz
                if k2 == 0 and not keep_k0:
                    continue
                if max(abs(kx), abs(ky), abs(kz)) >= thr:
                    continue
                if r2_of(k2):
                    cnt += 1
    return cnt


@pytest.mark.parametrize("n", [16, 32])
def test_spherical_mask_counts(n):
    """rhs.f90:74-78 removes |k|^2 >= (N/4)^2; state.f90:186-188 removes |k|^2 >= (N/3)^2 (the
    two rules differ, as written); counted by enumeration over the stored half spectrum."""
    o = Oracle(n, spherical=True)
    rhs_kept = int((~o.mask_array()).sum())
    assert rhs_kept == brute_unmasked(n, lambda k2: k2 < (n / 4.0) ** 2)
    ones = [np.ones(o.spec_shape, complex) for _ in range(3)]
    d = o.dealias(ones)
    assert int((d[0] != 0).sum()) == brute_unmasked(n, lambda k2: k2 < (n / 3.0) ** 2)


def test_spherical_single_modes():
    o = Oracle(32, spherical=True)
    # rhs rule (N/4)^2 = 64
    assert not o.masked(7, 3, 0) and o.masked(8, 0, 0) and o.masked(7, 4, 0)
    # dealias rule (N/3)^2 = 113.8: (10, 3, 0) -> 109 kept, (10, 4, 0) -> 116 removed
    f = [np.zeros(o.spec_shape, complex) for _ in range(3)]
    f[0][10, 3, 0] = 1.0
    f[0][10, 4, 0] = 1.0
    d = o.dealias(f)
    assert d[0][10, 3, 0] == 1.0 and d[0][10, 4, 0] == 0.0


def test_parity_mode_zero_removes_nyquist_modes():
    n = 16
    o = Oracle(n, parity_mode=False)
    m = o.mask_array()
    assert m[n // 2].all() and m[:, n // 2].all() and m[:, :, n // 2].all()
    kept = (~m).sum()
    # every other mode keeps its cubic-rule verdict
    m1 = Oracle(n).mask_array()
    m1[n // 2] = True
    m1[:, n // 2] = True
    m1[:, :, n // 2] = True
    assert np.array_equal(m, m1) and kept > 0


def test_linear_forcing_q_closed_form():
    """rhs_int_fact with Q (rhs.f90:10-14): the shear wave of test_shear_wave... grows at
    exp((Q - nu k^2) t)."""
    n, kf, nu, Q, dt, steps = 16, 2, 0.01, 0.0667, 0.01, 30
    o = Oracle(n, nu=nu, Q=Q)
    f = o.zeros_spec()
    f[0][0, kf, 0] = -0.5j * n ** 3
    f[0][0, n - kf, 0] = 0.5j * n ** 3
    g = o.heun(f, steps, dt)
    assert abs(g[0][0, kf, 0] / f[0][0, kf, 0] - np.exp((Q - nu * kf * kf) * steps * dt)) < 1e-13
