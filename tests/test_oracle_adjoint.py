"""Pins the oracle's adjoint pieces: the discrete RHS VJP (dot-product test against the
exact central difference of the quadratic RHS, and finite differences with band forcing),
the IF-Heun step VJP (finite differences), and the reference rhsAdj formula (rhsAdj.f90,
phatadj read as phat) against the discrete adjoint in the L2 (Hermitian-weighted) metric."""
import numpy as np
import pytest

from oracle_lib import Oracle


def inner(a, b):
    return float(sum(np.sum(np.real(np.conj(x) * y)) for x, y in zip(a, b)))


def rand_spec(o, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(o.spec_shape) + 1j * rng.standard_normal(o.spec_shape) for _ in range(3)]


def axpy(a, x, y):
    return [a * xi + yi for xi, yi in zip(x, y)]


@pytest.mark.parametrize("n,kolmo", [(16, 0), (32, 4)])
def test_rhs_vjp_dot_product(n, kolmo):
    o = Oracle(n, nu=0.03, kolmo_amp=1.0 if kolmo else 0.0, kolmo_k=kolmo)
    u = o.init_solenoidal(1, 0, 4.0, 0.5)
    v = rand_spec(o, 2)
    w = rand_spec(o, 3)
    # N is quadratic (+ a constant force): the central difference is exact for any eps
    jv = [(a - b) / 2.0 for a, b in zip(o.rhs(axpy(1.0, v, u)), o.rhs(axpy(-1.0, v, u)))]
    jtw = o.rhs_vjp(u, w)
    lhs, rhs = inner(jv, w), inner(v, jtw)
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1.0)


def test_rhs_vjp_band_forcing_fd():
    o = Oracle(16, nu=0.01, Pin=0.1, kcut=2.5, bandlim=True)
    u = o.init_solenoidal(4, 0, 2.0, 0.5)
    v = rand_spec(o, 5)
    w = rand_spec(o, 6)
    eps = 1e-6
    jv = [(a - b) / (2 * eps) for a, b in zip(o.rhs(axpy(eps, v, u)), o.rhs(axpy(-eps, v, u)))]
    lhs, rhs = inner(jv, w), inner(v, o.rhs_vjp(u, w))
    assert abs(lhs - rhs) < 1e-7 * abs(lhs)


def test_heun_vjp_fd():
    o = Oracle(16, nu=0.02, kolmo_amp=1.0, kolmo_k=2)
    u = o.init_solenoidal(7, 0, 3.0, 0.5)
    v = o.init_solenoidal(8, 0, 3.0, 0.5)
    w = rand_spec(o, 9)
    dt, eps = 0.01, 1e-5
    jv = [(a - b) / (2 * eps) for a, b in zip(o.heun(axpy(eps, v, u), 1, dt), o.heun(axpy(-eps, v, u), 1, dt))]
    lhs, rhs = inner(jv, w), inner(v, o.heun_vjp(u, w, dt))
    assert abs(lhs - rhs) < 1e-8 * abs(lhs)


def test_rhs_adj_equals_weighted_discrete_adjoint():
    """With the Nyquist modes removed (parity_mode = 0) and N = 32 (".ge." == ".gt."), the
    reference's continuous adjoint formula (phat fix) r = L u~ - P n_adj satisfies
        2 L u~ - r = P M DF^+(u~),  DF^+ = W^-1 J_F^T W,  u~ = H (L u + N(u)),
    i.e. the L2 adjoint of F = L + N restricted to retained solenoidal fields.  (The reference's L-term sign is therefore that of
    -DF^+ reflected: its linear part has the opposite sign of the descent direction.)"""
    n = 32
    o = Oracle(n, nu=0.02, parity_mode=False)
    u = o.init_solenoidal(3, 0, 4.0, 0.5)
    k = np.meshgrid(o.ktable(n)[: n // 2 + 1], o.ktable(n), o.ktable(n), indexing="ij")
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    L = -0.02 * k2
    H = 1.0 / (1.0 + k2)
    N = o.rhs(u)
    ut = [H * (L * a + b) for a, b in zip(u, N)]
    W = o.hermitian_weights()
    jt = o.rhs_vjp(u, [W * a for a in ut])
    # the discrete adjoint also has gradient (non-solenoidal) components; the continuous
    # formula is its projection onto retained solenoidal fields
    kk = np.where(k2 == 0, 1.0, k2)
    g = [b / W for b in jt]
    kg = (k[0] * g[0] + k[1] * g[1] + k[2] * g[2]) / kk
    m = o.mask_array()
    pg = [np.where(m, 0.0, g[i] - k[i] * kg) for i in range(3)]
    dfdag = [L * a + b for a, b in zip(ut, pg)]
    r = o.rhs_adj(u)
    expect = [2 * L * a - b for a, b in zip(ut, dfdag)]
    scale = max(np.abs(x).max() for x in expect)
    # the reference's ".gt." mask does not exclude k = 0, where it leaves -n_adj(0), the mean
    # of (grad u~ + grad u~^T) u; compare every other mode
    for x, y in zip(r, expect):
        x = x.copy()
        x[0, 0, 0] = 0.0
        assert np.abs(x - y).max() < 1e-10 * scale


def test_equilibrium_residual_gradient_fd():
    """J(u) = 1/2 sum W H |F(u)|^2 has W-metric gradient DF^+(u~) = 2 L u~ - r(u)."""
    n = 16
    o = Oracle(n, nu=0.05, parity_mode=False)
    u = o.init_solenoidal(5, 0, 3.0, 0.5)
    v = o.init_solenoidal(6, 0, 3.0, 0.5)
    k = np.meshgrid(o.ktable(n)[: n // 2 + 1], o.ktable(n), o.ktable(n), indexing="ij")
    k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
    L, H, W = -0.05 * k2, 1.0 / (1.0 + k2), o.hermitian_weights()

    def J(x):
        F = [L * a + b for a, b in zip(x, o.rhs(x))]
        return 0.5 * sum(np.sum(W * H * np.abs(f) ** 2) for f in F)

    eps = 1e-6
    fd = (J(axpy(eps, v, u)) - J(axpy(-eps, v, u))) / (2 * eps)
    N = o.rhs(u)
    ut = [H * (L * a + b) for a, b in zip(u, N)]
    r = o.rhs_adj(u)
    grad = [2 * L * a - b for a, b in zip(ut, r)]
    an = sum(np.sum(W * np.real(np.conj(g) * x)) for g, x in zip(grad, v))
    assert abs(fd - an) < 1e-6 * abs(fd)
