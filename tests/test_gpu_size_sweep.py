"""Every box size up to 128^3 in both precisions against the FP64 oracle: IF-Heun steps, the RHS
and step adjoints and rhsAdj.  The sizes run different kernel configurations (radix plans, tile
widths, lines per block; the plane pass up to 32^3), so each is checked on its own; FP32 at FP32
rounding (bar 2e-5), FP64 at 1e-12 (steps 1e-11)."""
import numpy as np
import pytest

from oracle_lib import Oracle, from_torch_layout, to_torch_layout

pytestmark = pytest.mark.gpu

SIZES = [(16, 64), (64, 64), (128, 64), (16, 32), (32, 32), (48, 32), (64, 32), (128, 32)]


def rel_l2(a, b):
    num = sum(np.linalg.norm((x - y).ravel()) ** 2 for x, y in zip(a, b))
    return float(np.sqrt(num / sum(np.linalg.norm(y.ravel()) ** 2 for y in b)))


def dev(f, prec):
    import torch
    t = torch.from_numpy(np.stack([to_torch_layout(a) for a in f])).cuda()
    return (t if prec == 64 else t.to(torch.complex64)).contiguous()


def host(t):
    import torch
    a = t.detach().cpu().to(torch.complex128).numpy()
    return [from_torch_layout(a[i]) for i in range(3)]


def rand_spec(o, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(o.spec_shape) + 1j * rng.standard_normal(o.spec_shape) for _ in range(3)]


def make(n, prec, **kw):
    from dnsbox_amd import Box
    return Box(n=n, precision=prec, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, **kw)


@pytest.mark.parametrize("n,prec", SIZES)
def test_heun_matches_oracle(gpu_lib, n, prec):
    dt = 0.004
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(90 + n, 0, 3.0, 0.5)
    ref = o.heun(f, 2, dt)
    with make(n, prec, dt=dt) as b:
        b.set_state(*f)
        b.rhs_int_fact(dt)
        b.step(2)
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"heun n={n} prec={prec}: rel_l2={err:.3e}", flush=True)
    assert err < (1e-11 if prec == 64 else 2e-5)


@pytest.mark.parametrize("n,prec", SIZES)
def test_adjoints_match_oracle(gpu_lib, n, prec):
    dt = 0.004
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    u = o.init_solenoidal(70 + n, 0, 3.0, 0.5)
    lam = rand_spec(o, n)
    refs = {"rhs_vjp": o.rhs_vjp(u, lam), "step_vjp": o.heun_vjp(u, lam, dt)}
    with make(n, prec, dt=dt) as b:
        out = b.empty_state()
        b.rhs_vjp_dev(dev(u, prec), dev(lam, prec), out)
        got = {"rhs_vjp": host(out)}
        b.step_vjp_dev(dev(u, prec), dev(lam, prec), out, dt)
        got["step_vjp"] = host(out)
    for k, ref in refs.items():
        err = rel_l2(got[k], ref)
        print(f"{k} n={n} prec={prec}: rel_l2={err:.3e}", flush=True)
        assert err < (1e-12 if prec == 64 else 2e-5), k


@pytest.mark.parametrize("n,prec", SIZES)
def test_rhs_adj_matches_oracle_all_sizes(gpu_lib, n, prec):
    o = Oracle(n, nu=0.02)
    u = o.init_solenoidal(50 + n, 0, 3.0, 0.5)
    ref = o.rhs_adj(u)
    from dnsbox_amd import Box
    with Box(n=n, precision=prec, nu=0.02, bandlim=0, adaptive_dt=0) as b:
        out = b.empty_state()
        b.rhs_adj_dev(dev(u, prec), out, 1)
        got = host(out)
    err = rel_l2(got, ref)
    print(f"rhs_adj n={n} prec={prec}: rel_l2={err:.3e}", flush=True)
    assert err < (1e-11 if prec == 64 else 2e-5)
