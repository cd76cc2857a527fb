"""Pins the oracle on non-cubic grids (Nx, Ny, Nz different, parameters.f90:12-14) before the GPU
path is compared with it there.  The transforms are pinned by numpy and a brute-force DFT
(test_oracle.py); here: an independent convective-form right-hand side with numpy.fft, the
invariants of the nonlinear term in the scaled wavenumbers, an exactly decaying shear wave,
Nspec and the Courant number from their formulas with hand-computed values, and the seeded
generator's solenoidality, Hermitian symmetry and energy.  A swapped axis length or an alpha on
the wrong axis fails one of them."""
import numpy as np
import pytest

from oracle_lib import Oracle, to_torch_layout

GRIDS = [((12, 16, 8), (1.0, 1.0, 1.0)), ((16, 8, 12), (1.0, 1.0, 0.5)), ((8, 12, 16), (2.0, 1.0, 0.5))]


def spec_k(o):
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


def make(shape, alpha, **kw):
    nx, ny, nz = shape
    return Oracle(nx, ny=ny, nz=nz, alpha=alpha, **kw)


def numpy_convective_rhs(o, f):
    """-P(u . grad u) with numpy.fft in the convective form, on the scaled wavenumbers."""
    kt = [to_torch_layout(k) for k in spec_k(o)]
    tl = [to_torch_layout(a) for a in f]
    u = [np.fft.irfftn(a, s=o.phys_shape, axes=(0, 1, 2)) for a in tl]
    grads = [[np.fft.irfftn(1j * kt[j] * tl[i], s=o.phys_shape, axes=(0, 1, 2)) for j in range(3)] for i in range(3)]
    ch = [np.fft.rfftn(sum(u[j] * grads[i][j] for j in range(3))) for i in range(3)]
    k2 = sum(k * k for k in kt)
    k2[k2 == 0] = 1
    kc = sum(kt[j] * ch[j] for j in range(3)) / k2
    return [np.transpose(-(ch[i] - kt[i] * kc), (2, 1, 0)) for i in range(3)]


@pytest.mark.parametrize("shape,alpha", GRIDS)
def test_rhs_rotational_equals_convective(shape, alpha):
    o = make(shape, alpha, nu=0.05)
    f = o.init_solenoidal(3, 0, 3.0, 0.5)
    got = o.rhs(f)
    ref = numpy_convective_rhs(o, f)
    keep = ~o.mask_array()
    keep[o.nx // 2, :, :] = False
    keep[:, o.ny // 2, :] = False
    keep[:, :, o.nz // 2] = False
    scale = max(np.abs(r[keep]).max() for r in ref)
    for g, r in zip(got, ref):
        assert np.abs(g[keep] - r[keep]).max() < 1e-12 * scale


@pytest.mark.parametrize("shape,alpha", GRIDS)
def test_rhs_invariants(shape, alpha):
    o = make(shape, alpha, nu=0.05)
    f = o.init_solenoidal(4, 0, 3.0, 0.5)
    nl = o.rhs(f)
    kx, ky, kz = spec_k(o)
    scale = max(np.abs(a).max() for a in nl)
    assert np.abs(kx * nl[0] + ky * nl[1] + kz * nl[2]).max() < 1e-12 * scale * max(shape)
    m = o.mask_array()
    assert all(np.all(a[m] == 0) for a in nl)
    w = hermitian_weights(o)
    dedt = sum(np.sum(w * np.real(np.conj(a) * b)) for a, b in zip(f, nl))
    mag = sum(np.sum(w * np.abs(a) * np.abs(b)) for a, b in zip(f, nl))
    assert abs(dedt) < 1e-12 * mag


def test_mask_per_axis():
    """rhs.f90:80-85: |k_d| >= (N_d/2)(2 alpha_d/3) per axis; at (12, 16, 8) the kept index ranges
    are |i| <= 3, 5, 2."""
    o = Oracle(12, ny=16, nz=8)
    assert not o.masked(3, 5, 2) and o.masked(4, 1, 1) and o.masked(1, 6, 1) and o.masked(1, 1, 3)
    assert not o.masked(1, 16 - 5, 8 - 2) and o.masked(1, 16 - 6, 1)


def test_nspec_per_axis():
    # variables.f90:117-122: int(sqrt(ex^2 + ey^2 + ez^2) / Deltak) - 1 with e_d = (N_d/2)(2 alpha_d/3)
    # (32, 48, 16): 10.667, 16, 5.333 -> sqrt(398.2) = 19.96 -> 18
    assert Oracle(32, ny=48, nz=16).nspec() == 18
    # (32, 32, 64), alpha_z = 1/2: 10.667, 10.667, 10.667 -> 18.48 -> 17
    assert Oracle(32, ny=32, nz=64, alpha=(1.0, 1.0, 0.5)).nspec() == 17
    # (256, 256, 128): 85.33, 85.33, 42.67 -> 128.0 -> 127
    assert Oracle(256, ny=256, nz=128).nspec() == 127


@pytest.mark.parametrize("d", [0, 1, 2])
def test_courant_per_axis(d):
    """io.f90:664-676: sum_d max|u_d| dt / (L_d / N_d).  A field with one unit-amplitude component
    (u_d = cos(x_e) along another axis e) gives dt N_d alpha_d / (2 pi)."""
    shape, alpha = (12, 16, 8), (1.0, 2.0, 0.5)
    o = make(shape, alpha)
    e = (d + 1) % 3
    phys = [np.zeros(o.phys_shape) for _ in range(3)]
    coords = np.meshgrid(*[np.arange(n) * 2 * np.pi / n for n in (o.nz, o.ny, o.nx)], indexing="ij")
    phys[d] = np.cos(coords[2 - e])
    f = [o.r2c(p) for p in phys]
    dt = 0.01
    assert abs(o.courant(f, dt) - dt * shape[d] * alpha[d] / (2 * np.pi)) < 1e-12


@pytest.mark.parametrize("shape,alpha", GRIDS)
def test_shear_wave_decays_exactly(shape, alpha):
    """u = sin(k y) e_x is a steady solution of the Euler terms (u . grad u = 0); with the
    integrating factor it decays as exp(-nu (alpha_y k)^2 t) in every step."""
    o = make(shape, alpha, nu=0.05)
    k = 2
    y = np.arange(o.ny) * 2 * np.pi / o.ny
    phys = [np.broadcast_to(np.sin(k * y)[None, :, None], o.phys_shape).copy(), np.zeros(o.phys_shape),
            np.zeros(o.phys_shape)]
    f = [o.r2c(p) for p in phys]
    dt, steps = 0.02, 5
    g = o.heun(f, steps, dt)
    decay = np.exp(-0.05 * (alpha[1] * k) ** 2 * dt * steps)
    assert max(np.abs(a - decay * b).max() for a, b in zip(g, f)) < 1e-12 * np.abs(f[0]).max()


@pytest.mark.parametrize("shape,alpha", GRIDS)
def test_init_solenoidal_on_grid(shape, alpha):
    o = make(shape, alpha)
    f = o.init_solenoidal(5, 0, 3.0, 0.5)
    kx, ky, kz = spec_k(o)
    assert np.abs(kx * f[0] + ky * f[1] + kz * f[2]).max() < 1e-13 * max(np.abs(a).max() for a in f)
    d, _ = o.diagnostics(f, 0.01)
    assert abs(d[0] - 0.5) < 1e-12
    m = o.nx * o.ny * o.nz
    back = [o.r2c(o.c2r(a) / m) for a in f]
    assert max(np.abs(a - b).max() for a, b in zip(f, back)) < 1e-13 * max(np.abs(a).max() for a in f)
    for a in f:
        assert np.all(a[o.mask_array()] == 0)
        assert np.all(a[o.nx // 2] == 0) and np.all(a[:, o.ny // 2] == 0) and np.all(a[:, :, o.nz // 2] == 0)
