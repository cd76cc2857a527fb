"""Initial fields and the analytic known-answer test of the reference, on a Box.

  shapiro(n, nu, t)        state_init_analytic (state.f90:287-307): the Shapiro/Beltrami
                           field, an exact solution decaying as exp(-3 nu t)
  init_shapiro(box, t)     that field as the box state (then dealias + project, as the
                           drivers do after loading a field, dnsboxheun.f90:51-55)
  shapiro_error(box, t)    state_check_error (state.f90:309-348): sum over the components of
                           max |u - u_exact| at time t
  init_yoffe(box, seed)    state_init_rand (state.f90:211-285): uniform random physical field,
                           dealias, project, then every shell rescaled to the spectrum
                           C1 k^C2 exp(-C3 k^C4) (Yoffe 2012 s. 2.2.5).  The reference draws
                           from an unseeded generator; here the draw is seeded.  With
                           reference_w=True the reference's w is reproduced as written: u is
                           drawn twice and w is left at -1/2 (state.f90:235-236)
  Box.init_random          the seeded solenoidal generator E(k) ~ k^4 exp(-2 (k/k0)^2) with
                           random phases (the Rosales-Meneveau form of python/dnsbox.py:259-281,
                           normalised to a given energy), run on the device

Physical arrays are (Nz, Ny, Nx), x fastest (the reference's u(Nx, Ny, Nz)); the grid starts
at -pi (variables.f90).  These are set-up routines, run once per run on the host.
"""
from __future__ import annotations

import numpy as np

YOFFE_C = (8.0e-2, 2.0, 8.2352309e-2, 2.0)   # state.f90:218-221 (the "S7" column)


def grid(n):
    """x_i = -pi + i 2 pi / n"""
    return (np.arange(n) - n // 2) * (2.0 * np.pi / n)


def shapiro(n, nu, t=0.0, ny=None, nz=None):
    """state_init_analytic (state.f90:287-307) at time t: (u, v, w), each of shape (nz, ny, n) on the
    2 pi box (ny, nz default n)."""
    f = np.sqrt(3.0)
    z, y, x = np.meshgrid(grid(nz or n), grid(ny or n), grid(n), indexing="ij")
    e = np.exp(-(f ** 2) * t * nu)
    u = -0.5 * (f * np.cos(x) * np.sin(y) * np.sin(z) + np.sin(x) * np.cos(y) * np.cos(z)) * e
    v = 0.5 * (f * np.sin(x) * np.cos(y) * np.sin(z) - np.cos(x) * np.sin(y) * np.cos(z)) * e
    w = 1.0 * np.cos(x) * np.cos(y) * np.sin(z) * e
    return u, v, w


def init_shapiro(box, t=0.0):
    box.set_state_physical(*shapiro(box.nx, box.params.nu, t, box.ny, box.nz))
    box.state_dealias()
    box.state_project()


def shapiro_error(box, t):
    """state_check_error (state.f90:309-348): sum_i max |u_i - u_i,exact(t)|."""
    ex = shapiro(box.nx, box.params.nu, t, box.ny, box.nz)
    got = box.get_state_physical()
    return float(sum(np.max(np.abs(a - b)) for a, b in zip(got, ex)))


def _ktable(n):
    """variables.f90:226-262: wavenumbers with the Nyquist entry set to 0."""
    k = np.arange(n, dtype=np.float64)
    k = np.where(k < n // 2, k, np.where(k == n // 2, 0.0, k - n))
    return k


def shell_index(absk, deltak=1.0):
    """state.f90:246-251 / :363-367 as written: int(|k|/dk), plus one when |k| is at or past
    real(int(|k|/dk)) + dk/2 (real() is single precision there)."""
    iq = np.floor(absk / deltak).astype(np.int64)
    half = iq.astype(np.float32).astype(np.float64) + 0.5 * deltak
    return np.where(absk < half, iq, iq + 1)


def yoffe_spectrum(k, c=YOFFE_C):
    c1, c2, c3, c4 = c
    return c1 * k ** c2 * np.exp(-c3 * k ** c4)


def init_yoffe(box, seed=0, c=YOFFE_C, reference_w=True):
    """state_init_rand (state.f90:211-285) on the box, seeded."""
    shape = box.phys_shape
    rng = np.random.default_rng(seed)
    u = rng.random(shape) - 0.5
    v = rng.random(shape) - 0.5
    if reference_w:
        rng.random(shape)                  # the reference draws u a second time ...
        w = np.full(shape, -0.5)            # ... and shifts an untouched w by -1/2
    else:
        w = rng.random(shape) - 0.5
    box.set_state_physical(u, v, w)
    box.state_dealias()
    box.state_project()
    _, espec = box.diagnostics()
    deltak = box.params.deltak or 1.0
    nspec = len(espec)
    # reference layout uhat(Nz, Ny, Nx/2+1), z fastest: arrays indexed [kx][ky][kz]
    f = box.get_state()
    p = box.params
    ax, ay, az = (float(np.float32(a)) if a > 0 else 1.0 for a in (p.alpha_x, p.alpha_y, p.alpha_z))
    kx = (_ktable(box.nx)[: box.nx // 2 + 1] * ax).reshape(-1, 1, 1)
    ky = (_ktable(box.ny) * ay).reshape(1, -1, 1)
    kz = (_ktable(box.nz) * az).reshape(1, 1, -1)
    absk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    nk = shell_index(absk, deltak)
    keep = (nk > 0) & (nk < nspec)
    es = np.zeros(nspec + 1)
    es[1:] = espec
    target = yoffe_spectrum(nk * deltak, c)
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.where(keep & (es[np.minimum(nk, nspec)] > 0), np.sqrt(target / es[np.minimum(nk, nspec)]), 0.0)
    box.set_state(*[a * scale for a in f])
    return scale
