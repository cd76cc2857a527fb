"""Pins the oracle's adjoint-descent step (orc_adj_descent, dnsboxAdj.f90:76-126): Heun in
fictitious time tau on rhsAdj, then state_project.

* Small amplitude: rhsAdj(u) = lin_sign L H L u + O(|u|^2) (L = -nu |k|^2 + Q, H = 1/(1+|k|^2)),
  so one step multiplies each mode by 1 + lam dtau + (lam dtau)^2 / 2, lam = lin_sign L^2 H:
  pins both half weights of the Heun update, H and the sign.
* Second order in dtau at full amplitude.
* With lin_sign = -1 (the descent direction) the residual J = 1/2 <F, H F>, F = L u + N(u),
  decreases; with the reference's sign (+1) it grows for a small field (the reference's linear
  term has the sign opposite to -DF^+).
"""
import numpy as np
import pytest

from oracle_lib import Oracle


def spec_k2(o):
    kx = o.ktable(o.nx)[: o.nx // 2 + 1]
    ky = o.ktable(o.ny)
    kz = o.ktable(o.nz)
    x, y, z = np.meshgrid(kx, ky, kz, indexing="ij")
    return x * x + y * y + z * z


def _err(a, b):
    return np.sqrt(sum(np.sum(np.abs(x - y) ** 2) for x, y in zip(a, b)))


def _norm(a):
    return np.sqrt(sum(np.sum(np.abs(x) ** 2) for x in a))


@pytest.mark.parametrize("lin_sign", [1, -1])
def test_small_amplitude_closed_form(lin_sign):
    n, nu, Q, dtau, eps = 16, 0.2, 0.02, 1.0, 1e-7
    o = Oracle(n, nu=nu, Q=Q)
    f = [eps * a for a in o.init_solenoidal(2, 0, 3.0, 0.5)]
    g = o.adj_descent(f, dtau, 1, lin_sign)
    k2 = spec_k2(o)
    L = -nu * k2 + Q
    lam = lin_sign * L * L / (1.0 + k2)
    fac = 1.0 + lam * dtau + 0.5 * (lam * dtau) ** 2
    expect = [fac * a for a in f]
    assert _err(g, expect) < 1e-5 * _norm(expect)
    # a dropped half weight or H would be far off
    assert _err(g, [(1.0 + lam * dtau) * a for a in f]) > 1e-3 * _norm(expect)


def test_second_order_in_dtau():
    o = Oracle(16, nu=0.05)
    f = o.init_solenoidal(4, 0, 3.0, 0.1)
    T = 0.2
    fine = o.adj_descent(f, T / 64, 64, -1)
    errs = [_err(o.adj_descent(f, T / m, m, -1), fine) for m in (4, 8, 16)]
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 4.8 and 3.3 < r2 < 4.8, (errs, r1, r2)


def residual(o, f, nu, Q):
    """J = 1/2 sum W H |L u + N(u)|^2 / N^6 (the H^-1 norm of the equilibrium residual)."""
    k2 = spec_k2(o)
    L = -nu * k2 + Q
    nl = o.rhs(f)
    w = o.hermitian_weights()
    h = 1.0 / (1.0 + k2)
    return 0.5 * sum(np.sum(w * h * np.abs(L * a + b) ** 2) for a, b in zip(f, nl)) / float(o.nx) ** 6


def test_descent_direction_lowers_the_residual():
    n, nu = 16, 0.05
    o = Oracle(n, nu=nu)
    f = o.init_solenoidal(6, 0, 3.0, 0.5)
    j0 = residual(o, f, nu, 0.0)
    js = [residual(o, o.adj_descent(f, d, 1, -1), nu, 0.0) for d in (0.02, 0.05)]
    assert all(j < j0 for j in js), (j0, js)
    small = [1e-3 * a for a in f]
    js0 = residual(o, small, nu, 0.0)
    assert residual(o, o.adj_descent(small, 0.05, 1, 1), nu, 0.0) > js0


def test_one_step_is_the_two_stage_composition():
    o = Oracle(16, nu=0.05)
    f = o.init_solenoidal(7, 0, 3.0, 0.5)
    d = 0.04
    g0 = o.rhs_adj(f)
    us = [a + d * b for a, b in zip(f, g0)]
    g1 = o.rhs_adj(us)
    u1 = o.project([a + d * b * 0.5 + d * c * 0.5 for a, b, c in zip(f, g0, g1)])
    assert _err(o.adj_descent(f, d, 1, 1), u1) < 1e-14 * _norm(u1)
