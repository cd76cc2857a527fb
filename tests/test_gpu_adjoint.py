"""GPU adjoint kernels vs the CPU oracle, and the differentiable T-step map vs central
finite differences (through the C-ABI, FP64)."""
import numpy as np
import pytest

from oracle_lib import Oracle, from_torch_layout, to_torch_layout

pytestmark = pytest.mark.gpu


def dev(f):
    import torch
    return torch.from_numpy(np.stack([to_torch_layout(a) for a in f])).cuda().contiguous()


def host(t):
    a = t.detach().cpu().numpy()
    return [from_torch_layout(a[i]) for i in range(3)]


def rel_l2(a, b):
    num = sum(np.linalg.norm((x - y).ravel()) ** 2 for x, y in zip(a, b))
    return float(np.sqrt(num / sum(np.linalg.norm(y.ravel()) ** 2 for y in b)))


def make(n, **kw):
    from dnsbox_amd import Box
    return Box(n=n, precision=kw.pop("precision", 64), **kw)


def rand_spec(o, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(o.spec_shape) + 1j * rng.standard_normal(o.spec_shape) for _ in range(3)]


@pytest.mark.parametrize("n", [32, 48])
def test_rhs_vjp_matches_oracle(gpu_lib, n):
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    u = o.init_solenoidal(1, 0, 4.0, 0.5)
    lam = rand_spec(o, 2)
    ref = o.rhs_vjp(u, lam)
    with make(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0) as b:
        out = b.empty_state()
        b.rhs_vjp_dev(dev(u), dev(lam), out)
        got = host(out)
    err = rel_l2(got, ref)
    print(f"rhs_vjp n={n} rel_l2={err:.3e}", flush=True)
    assert err < 1e-12


def test_rhs_vjp_band_matches_oracle(gpu_lib):
    n = 32
    o = Oracle(n, nu=0.01, Pin=0.1, kcut=2.5, bandlim=True)
    u = o.init_solenoidal(3, 0, 2.0, 0.5)
    lam = rand_spec(o, 4)
    ref = o.rhs_vjp(u, lam)
    with make(n, nu=0.01, Pin=0.1, kcut=2.5, bandlim=1, adaptive_dt=0) as b:
        out = b.empty_state()
        b.rhs_vjp_dev(dev(u), dev(lam), out)
        got = host(out)
    err = rel_l2(got, ref)
    print(f"rhs_vjp band rel_l2={err:.3e}", flush=True)
    assert err < 1e-12


def test_step_vjp_matches_oracle(gpu_lib):
    n = 32
    dt = 0.005
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    u = o.init_solenoidal(5, 0, 4.0, 0.5)
    lam = rand_spec(o, 6)
    ref = o.heun_vjp(u, lam, dt)
    with make(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        out = b.empty_state()
        b.step_vjp_dev(dev(u), dev(lam), out, dt)
        got = host(out)
    err = rel_l2(got, ref)
    print(f"step_vjp rel_l2={err:.3e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("recorded", [False, True])
def test_step_vjp_band_matches_oracle(gpu_lib, recorded):
    """The step adjoint with band forcing (the reference default, parameters.f90:52): the band
    terms of the epilogue at lam_scale = dt/2 (stage 1) and 1 (stage 2), from the recomputed or
    the supplied predicted state."""
    n, dt = 32, 0.004
    o = Oracle(n, nu=0.01, Pin=0.1, kcut=2.5, bandlim=True)
    u = o.init_solenoidal(3, 0, 2.0, 0.5)
    lam = rand_spec(o, 8)
    ref = o.heun_vjp(u, lam, dt)
    with make(n, nu=0.01, Pin=0.1, kcut=2.5, bandlim=1, adaptive_dt=0, dt=dt) as b:
        out = b.empty_state()
        if recorded:
            import torch
            st = dev(u)
            traj = torch.empty((1,) + tuple(st.shape), dtype=st.dtype, device=st.device)
            pred = torch.empty_like(traj)
            b.rhs_int_fact(dt)
            b.step_dev_record(st.clone(), 1, traj, pred)
            b.step_vjp_pred_dev(dev(u), pred[0], dev(lam), out, dt)
        else:
            b.step_vjp_dev(dev(u), dev(lam), out, dt)
        got = host(out)
    err = rel_l2(got, ref)
    print(f"step_vjp band recorded={recorded} rel_l2={err:.3e}", flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("n,sign", [(32, 1), (32, -1), (48, 1)])
def test_rhs_adj_matches_oracle(gpu_lib, n, sign):
    o = Oracle(n, nu=0.02)
    u = o.init_solenoidal(7, 0, 4.0, 0.5)
    ref = o.rhs_adj(u)
    if sign < 0:
        k = np.meshgrid(o.ktable(n)[: n // 2 + 1], o.ktable(n), o.ktable(n), indexing="ij")
        k2 = k[0] ** 2 + k[1] ** 2 + k[2] ** 2
        L, H = -0.02 * k2, 1.0 / (1.0 + k2)
        ut = [H * (L * a + b) for a, b in zip(u, Oracle(n, nu=0.02).rhs(u))]
        ref = [r - 2 * L * a for r, a in zip(ref, ut)]
    with make(n, nu=0.02, bandlim=0, adaptive_dt=0) as b:
        out = b.empty_state()
        b.rhs_adj_dev(dev(u), out, sign)
        got = host(out)
    m = np.ones(ref[0].shape, bool)
    if sign < 0:  # u~ above uses the rotational RHS: compare off the Nyquist planes and k = 0
        m[n // 2] = False
        m[:, n // 2] = False
        m[:, :, n // 2] = False
        m[0, 0, 0] = False
    num = sum(np.linalg.norm((g - r)[m]) ** 2 for g, r in zip(got, ref))
    den = sum(np.linalg.norm(r[m]) ** 2 for r in ref)
    err = float(np.sqrt(num / den))
    print(f"rhs_adj n={n} sign={sign} rel_l2={err:.3e}", flush=True)
    assert err < 1e-11


def test_step_dev_matches_internal_state(gpu_lib):
    n = 32
    kw = dict(nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.005)
    with make(n, **kw) as a:
        a.init_random(3, 0, 4.0, 0.5)
        f = a.get_state()
        t = dev(f)
        a.rhs_int_fact(0.005)
        a.step_dev(t, 10)
        a.set_state(*f)
        a.set_time(0.0, 0.005)
        a.step(10)
        ref = a.get_state()
    assert rel_l2(host(t), ref) < 1e-14


def test_step_dev_prunes_only_dealiased_buffers(gpu_lib):
    """step_dev checks the caller's buffer on the device: a field with energy outside the
    kept cube is stepped with the full passes (matches the oracle, masked modes included)
    even when the box's own state is dealiased; step_dev_record stores the trajectory."""
    import torch
    n, dt = 32, 0.005
    o = Oracle(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(0, 0, 4.0, 0.5)
    g = [x.copy() for x in f]
    g[0][12, 3, 5] = 1e-3 + 2e-3j   # kx = 12 > (32 - 1) // 3
    with make(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.init_random(1, 0, 4.0, 0.5)
        assert b.pruning()
        b.rhs_int_fact(dt)
        for field, name in ((f, "dealiased"), (g, "masked-mode energy")):
            t = dev(field)
            b.step_dev(t, 3)
            err = rel_l2(host(t), o.heun(field, 3, dt))
            print(f"step_dev {name}: rel_l2 vs oracle {err:.3e}", flush=True)
            assert err <= 1e-12
        t0 = dev(f)
        traj = torch.empty((4,) + tuple(t0.shape), dtype=t0.dtype, device=t0.device)
        preds = torch.full_like(traj, 7.0)   # stale values outside the kept cube must not survive
        u = t0.clone()
        b.step_dev_record(u, 4, traj, preds)
        ref = t0.clone()
        lam = dev(o.init_solenoidal(5, 0, 4.0, 0.5))
        for j in range(4):
            prev = ref.clone()
            b.step_dev(ref, 1)
            e = rel_l2(host(traj[j]), host(ref))
            # the recorded predicted state gives the step adjoint of the recomputed one
            o1, o2 = torch.empty_like(lam), torch.empty_like(lam)
            b.step_vjp_dev(prev, lam, o1, dt)
            b.step_vjp_pred_dev(prev, preds[j], lam, o2, dt)
            ev = rel_l2(host(o2), host(o1))
            print(f"record step {j + 1}: state rel_l2 {e:.3e}, step-VJP with recorded u* rel_l2 {ev:.3e}",
                  flush=True)
            assert e <= 1e-14 and ev <= 1e-14
        assert torch.equal(u, traj[3])


@pytest.mark.parametrize("n,T,k,record", [(32, 20, 5, "checkpoint"), (64, 200, 20, "auto")])
def test_tstep_gradient_vs_finite_differences(gpu_lib, n, T, k, record):
    """d/de J(u + e v) by central differences (e = 1e-6) along seeded solenoidal directions
    vs <grad J, v> from the adjoint (checkpointed, or with the trajectory recorded, which "auto"
    picks at 64^3, T = 200); tolerance 1e-6 relative."""
    import torch
    from dnsbox_amd.adjoint import residual
    dt = 0.005
    nu = 0.02
    o = Oracle(n, nu=nu, kolmo_amp=1.0, kolmo_k=4)
    ndir = 8
    with make(n, nu=nu, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        u0 = dev(o.init_solenoidal(11, 0, 4.0, 0.5))
        x = u0.clone().requires_grad_(True)
        J = residual(b, x, T, dt, k, record)
        (g,) = torch.autograd.grad(J, x)
        worst = 0.0
        for d in range(ndir):
            v = dev(o.init_solenoidal(100 + d, 0, 4.0, 0.5)) * 1.0
            eps = 1e-6
            with torch.no_grad():
                jp = float(residual(b, u0 + eps * v, T, dt, k))
                jm = float(residual(b, u0 - eps * v, T, dt, k))
            fd = (jp - jm) / (2 * eps)
            an = float((g.conj() * v).real.sum())
            rel = abs(fd - an) / abs(fd)
            worst = max(worst, rel)
            print(f"T-step grad n={n} T={T} dir={d}: fd={fd:.10e} adj={an:.10e} rel={rel:.2e}", flush=True)
        assert worst < 1e-6


@pytest.mark.parametrize("n,T,k", [(32, 40, 7), (64, 60, 20)])
def test_overlapped_backward_equals_one_stream_loop(gpu_lib, n, T, k):
    """The checkpointed backward (segments recomputed by the twin box on a second stream while the
    box runs the step adjoints of the next segment) gives the gradient of a plain loop that
    recomputes every state with single steps and back-propagates in one stream, to rounding
    (single steps restart the z-inverse from the state; a race would show as an O(1) error)."""
    import torch
    from dnsbox_amd.adjoint import step_map
    dt = 0.005
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    with make(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        u0 = dev(o.init_solenoidal(3, 0, 4.0, 0.5))
        g = dev(rand_spec(o, 4))
        x = u0.clone().requires_grad_(True)
        (got,) = torch.autograd.grad(step_map(b, x, T, dt, k, record="checkpoint"), x, g)
        # the same backward, one stream, no twin
        b.rhs_int_fact(dt)
        ckpts, u = [], u0.clone()
        for t in range(0, T, k):
            ckpts.append(u.clone())
            b.step_dev(u, min(k, T - t))
        lam, out = g.clone(), torch.empty_like(g)
        for seg in reversed(range(len(ckpts))):
            m = min(k, T - seg * k)
            states = [ckpts[seg]]
            for _ in range(m - 1):
                v = states[-1].clone()
                b.step_dev(v, 1)
                states.append(v)
            for j in reversed(range(m)):
                b.step_vjp_dev(states[j], lam, out, dt)
                lam, out = out, lam
        torch.cuda.synchronize()
        err = rel_l2(host(got), host(lam))
        print(f"overlapped backward n={n} T={T} k={k}: rel_l2 vs one-stream loop {err:.3e}", flush=True)
        assert err <= 1e-14


@pytest.mark.parametrize("n", [32, 64])
def test_recorded_trajectory_backward_equals_checkpointed(gpu_lib, n):
    """The backward through the trajectory recorded by the forward (record="all": no segment is
    recomputed) gives the checkpointed backward's gradient to rounding (the checkpointed path
    restarts the step pipeline at every segment, with an unfused z-inverse from the stored state)."""
    import torch
    from dnsbox_amd.adjoint import step_map
    dt, T, k = 0.005, 45, 10
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    with make(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        u0 = dev(o.init_solenoidal(8, 0, 4.0, 0.5))
        g = dev(rand_spec(o, 9))
        grads = {}
        for record in ("all", "checkpoint"):
            x = u0.clone().requires_grad_(True)
            (grads[record],) = torch.autograd.grad(step_map(b, x, T, dt, k, record=record), x, g)
        torch.cuda.synchronize()
    err = rel_l2(host(grads["all"]), host(grads["checkpoint"]))
    print(f"recorded vs checkpointed backward n={n}: rel_l2 {err:.3e}", flush=True)
    assert err <= 1e-13


@pytest.mark.parametrize("n", [32, 64])
def test_vjp_cube_guard_reports_a_state_outside_the_cube(gpu_lib, n):
    """set_vjp_cube(True) is checked on the device: differentiating a state with energy outside
    the kept cube under it (as u, or as a supplied u*) makes set_vjp_cube(False) raise; confined
    states pass, and a new guarantee starts from a cleared flag."""
    import torch
    from dnsbox_amd._lib import DnsboxError
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    dt = 0.005
    with make(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.rhs_int_fact(dt)
        u = dev(o.init_solenoidal(5, 0, 4.0, 0.5))
        ustar = u.clone()
        b.step_dev(u, 1)
        b.step_dev(ustar, 1)
        assert b.last_step_pruned()
        lam = dev(rand_spec(o, 6))
        w = dev(rand_spec(o, 7))   # energy on every mode
        out = torch.empty_like(u)
        b.set_vjp_cube(True)
        b.step_vjp_dev(u, lam, out, dt)
        b.step_vjp_pred_dev(u, ustar, lam, out, dt)
        b.rhs_vjp_dev(u, lam, out)
        b.set_vjp_cube(False)      # confined states: no error
        for bad in ("u", "ustar"):
            b.set_vjp_cube(True)
            if bad == "u":
                b.rhs_vjp_dev(w, lam, out)
            else:
                b.step_vjp_pred_dev(u, w, lam, out, dt)
            with pytest.raises(DnsboxError, match="outside"):
                b.set_vjp_cube(False)
        b.set_vjp_cube(True)
        b.step_vjp_dev(u, lam, out, dt)
        b.set_vjp_cube(False)
    print(f"vjp cube guard n={n}: confined states pass, u or u* outside the cube raise", flush=True)


@pytest.mark.parametrize("n,prec", [(32, 64), (64, 64), (48, 32)])
def test_vjp_cube_matches_full_passes(gpu_lib, n, prec):
    """With set_vjp_cube (linearisation states confined to the kept cube) the RHS and step
    adjoints read only the cube in their first passes and agree with the full passes to rounding;
    last_step_pruned reports the device check of the last step on a caller's buffer."""
    import torch
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    dt = 0.005
    dtype = torch.complex128 if prec == 64 else torch.complex64
    tol = 1e-13 if prec == 64 else 1e-6
    with make(n, precision=prec, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.rhs_int_fact(dt)
        u = dev(o.init_solenoidal(5, 0, 4.0, 0.5)).to(dtype)
        b.step_dev(u, 2)
        assert b.last_step_pruned()
        lam = dev(rand_spec(o, 6)).to(dtype)
        outs = {}
        for cube in (False, True):
            b.set_vjp_cube(cube)
            r = torch.empty_like(u)
            st = torch.empty_like(u)
            b.rhs_vjp_dev(u, lam, r)
            b.step_vjp_dev(u, lam, st, dt)
            torch.cuda.synchronize()
            outs[cube] = (host(r), host(st))
        b.set_vjp_cube(False)
        e_rhs = rel_l2(outs[True][0], outs[False][0])
        e_step = rel_l2(outs[True][1], outs[False][1])
        print(f"vjp cube n={n} prec={prec}: RHS-VJP rel {e_rhs:.2e}, step-VJP rel {e_step:.2e}", flush=True)
        assert e_rhs <= tol and e_step <= tol
        w = dev(rand_spec(o, 7)).to(dtype)   # energy on every mode
        b.step_dev(w, 1)
        assert not b.last_step_pruned()
