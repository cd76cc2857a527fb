"""The plane pass (the y-inverse, x and y-forward passes of one z-plane in one block, run by the
boxes up to 32^3) against the oracle: IF-Heun steps, pruned to the 2/3-rule cube and not, and
step adjoints with the linearisation state confined to the cube and not; FP32 at FP32 rounding.
Larger boxes run the three passes (checked by the byte counts of the step's passes)."""
import numpy as np
import pytest

from oracle_lib import Oracle, from_torch_layout, to_torch_layout

pytestmark = pytest.mark.gpu


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


# FP32 transforms of O(1) fields: a few FP32 ulp per transform, grown over a few steps
TOL = {64: 1e-11, 32: 2e-5}


@pytest.mark.parametrize("pruned", [True, False])
@pytest.mark.parametrize("n,prec", [(16, 64), (32, 64), (48, 64), (16, 32), (32, 32), (48, 32), (64, 32)])
def test_plane_heun_matches_oracle(gpu_lib, n, prec, pruned):
    from dnsbox_amd import Box
    dt, steps = 0.004, 6
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(40 + n, 0, 3.0, 0.5)
    ref = o.heun(f, steps, dt)
    with Box(n=n, precision=prec, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.set_pruning(pruned)
        b.set_state(*f)
        assert b.pruning() == pruned
        b.rhs_int_fact(dt)
        b.step(steps)
        nb = b.pass_bytes()
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"plane heun n={n} prec={prec} pruned={pruned}: rel_l2={err:.3e} "
          f"(middle passes: {'one plane launch' if nb['yinv6'] == 0 else 'three launches'})", flush=True)
    assert (nb["yinv6"] == 0 and nb["yfwd3"] == 0) == (n <= 32)   # the plane pass up to 32^3
    assert err < TOL[prec]


@pytest.mark.parametrize("cube", [True, False])
@pytest.mark.parametrize("n,prec", [(16, 64), (32, 64), (16, 32), (32, 32), (48, 32), (64, 32)])
def test_plane_step_vjp_matches_oracle(gpu_lib, n, prec, cube):
    from dnsbox_amd import Box
    dt = 0.004
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    u = o.init_solenoidal(60 + n, 0, 3.0, 0.5)
    rng = np.random.default_rng(n)
    lam = [rng.standard_normal(o.spec_shape) + 1j * rng.standard_normal(o.spec_shape) for _ in range(3)]
    ref = o.heun_vjp(u, lam, dt)
    with Box(n=n, precision=prec, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        out = b.empty_state()
        b.set_vjp_cube(cube)
        b.step_vjp_dev(dev(u, prec), dev(lam, prec), out, dt)
        b.set_vjp_cube(False)
        got = host(out)
    err = rel_l2(got, ref)
    print(f"plane step_vjp n={n} prec={prec} cube={cube}: rel_l2={err:.3e}", flush=True)
    assert err < (1e-12 if prec == 64 else TOL[32])
