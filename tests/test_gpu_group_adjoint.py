"""Batched (group) step and step adjoint on caller buffers vs the single-box entry points,
and the batched checkpointed T-step gradient vs per-box gradients and finite differences."""
import numpy as np
import pytest

from oracle_lib import Oracle, from_torch_layout, to_torch_layout

pytestmark = pytest.mark.gpu

SIZES = [32, 48, 64, 16, 32]
NUS = [0.02, 0.03, 0.015, 0.05, 0.04]
DT = 0.005


def dev(f, dtype):
    import torch
    return torch.from_numpy(np.stack([to_torch_layout(a) for a in f])).to(dtype).cuda().contiguous()


def host(t):
    a = t.detach().cpu().to(__import__("torch").complex128).numpy()
    return [from_torch_layout(a[i]) for i in range(3)]


def rel_l2(a, b):
    num = sum(np.linalg.norm((x - y).ravel()) ** 2 for x, y in zip(a, b))
    return float(np.sqrt(num / sum(np.linalg.norm(y.ravel()) ** 2 for y in b)))


def kw(n, nu, prec):
    return dict(n=n, precision=prec, nu=nu, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=DT)


def fields(prec, seed0):
    import torch
    dtype = torch.complex128 if prec == 64 else torch.complex64
    us = []
    for i, (n, nu) in enumerate(zip(SIZES, NUS)):
        o = Oracle(n, nu=nu, kolmo_amp=1.0, kolmo_k=4)
        us.append(dev(o.init_solenoidal(seed0 + i, 0, 2.0 + i, 0.5), dtype))
    return us


def rand_like(us, seed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    out = []
    for u in us:
        re = torch.randn(u.shape, generator=g, dtype=torch.float64)
        im = torch.randn(u.shape, generator=g, dtype=torch.float64)
        out.append(torch.complex(re, im).to(u.dtype).cuda().contiguous())
    return out


@pytest.mark.parametrize("prec", [64, 32])
def test_group_step_dev_and_vjp_match_single_boxes(gpu_lib, prec):
    import torch
    from dnsbox_amd import Box, BoxGroup
    tol = 1e-14 if prec == 64 else 1e-6
    us = fields(prec, 40)
    lams = rand_like(us, 7)
    with BoxGroup([kw(n, nu, prec) for n, nu in zip(SIZES, NUS)]) as g:
        for b in g.boxes:
            b.rhs_int_fact(DT)
        st = [u.clone() for u in us]
        g.step_dev(st, 3)
        outs = [torch.empty_like(u) for u in us]
        g.step_vjp_dev(us, lams, outs, [DT] * len(us))
        torch.cuda.synchronize()
    for i, (n, nu) in enumerate(zip(SIZES, NUS)):
        with Box(**kw(n, nu, prec)) as b:
            b.rhs_int_fact(DT)
            ref = us[i].clone()
            b.step_dev(ref, 3)
            o = torch.empty_like(us[i])
            b.step_vjp_dev(us[i], lams[i], o, DT)
            torch.cuda.synchronize()
        e1 = rel_l2(host(st[i]), host(ref))
        e2 = rel_l2(host(outs[i]), host(o))
        print(f"group box {i} n={n} prec={prec}: step_dev rel={e1:.2e} step_vjp rel={e2:.2e}", flush=True)
        assert e1 <= tol and e2 <= tol


def test_group_tstep_gradient_matches_single_and_fd(gpu_lib):
    """Batched checkpointed gradient of J_i = 1/2 <|Phi_T(u_i) - u_i|^2> equals the single-box
    one per box, and matches central differences (eps = 1e-6) to 1e-6 relative (FP64)."""
    import torch
    from dnsbox_amd import Box, BoxGroup
    from dnsbox_amd.adjoint import group_residual_grads, group_residuals, residual
    T, k = 30, 7
    us = fields(64, 60)
    with BoxGroup([kw(n, nu, 64) for n, nu in zip(SIZES, NUS)]) as g:
        dts = [DT] * len(us)
        J, gs = group_residual_grads(g, us, T, dts, k)
        vs = [0.1 * v for v in fields(64, 80)]
        eps = 1e-6
        with torch.no_grad():
            jp = group_residuals(g, [u + eps * v for u, v in zip(us, vs)], T, dts, k).tolist()
            jm = group_residuals(g, [u - eps * v for u, v in zip(us, vs)], T, dts, k).tolist()
    worst_fd = 0.0
    for i, (n, nu) in enumerate(zip(SIZES, NUS)):
        fd = (jp[i] - jm[i]) / (2 * eps)
        an = float((gs[i].conj() * vs[i]).real.sum())
        rel = abs(fd - an) / abs(fd)
        worst_fd = max(worst_fd, rel)
        with Box(**kw(n, nu, 64)) as b:
            x = us[i].clone().requires_grad_(True)
            Ji = residual(b, x, T, DT, k)
            (gi,) = torch.autograd.grad(Ji, x)
        eg = rel_l2(host(gs[i]), host(gi))
        print(f"group grad box {i} n={n}: J={float(J[i]):.6e} single={Ji.item():.6e} grad rel={eg:.2e} "
              f"fd={fd:.10e} adj={an:.10e} rel={rel:.2e}", flush=True)
        assert abs(float(J[i]) - Ji.item()) <= 1e-13 * abs(Ji.item())
        assert eg < 1e-13
    assert worst_fd < 1e-6


def test_group_adjoint_descent_decreases_every_residual(gpu_lib):
    from dnsbox_amd import BoxGroup
    from dnsbox_amd.adjoint import group_adjoint_descent
    T, k = 20, 5
    us = fields(32, 90)
    with BoxGroup([kw(n, nu, 32) for n, nu in zip(SIZES, NUS)]) as g:
        _, hist = group_adjoint_descent(g, us, T, [DT] * len(us), k, iters=4)
    first, last = hist[0], hist[-1]
    for i in range(len(us)):
        print(f"descent box {i}: J {first[i]:.4e} -> {last[i]:.4e}", flush=True)
        assert last[i] < first[i]


@pytest.mark.parametrize("prec", [64, 32])
def test_group_record_and_vjp_with_recorded_prediction(gpu_lib, prec):
    """dnsbox_group_step_dev_record stores every box's state after each step (equal to single
    group steps) and each step's predicted state u*; the step adjoint given that u* equals the
    one that recomputes it."""
    import torch
    from dnsbox_amd import BoxGroup
    tol = 1e-14 if prec == 64 else 1e-6
    nst = 4
    us = fields(prec, 50)
    lams = rand_like(us, 9)
    with BoxGroup([kw(n, nu, prec) for n, nu in zip(SIZES, NUS)]) as g:
        for b in g.boxes:
            b.rhs_int_fact(DT)
        trajs = [torch.empty((nst,) + tuple(u.shape), dtype=u.dtype, device=u.device) for u in us]
        preds = [torch.full_like(t, 7.0) for t in trajs]   # every slot must be written
        st = [u.clone() for u in us]
        g.step_dev_record(st, nst, trajs, preds)
        ref = [u.clone() for u in us]
        dts = [DT] * len(us)
        worst_s = worst_v = 0.0
        for j in range(nst):
            g.step_dev(ref, 1)
            uj = us if j == 0 else [t[j - 1] for t in trajs]
            a = [torch.empty_like(u) for u in us]
            b = [torch.empty_like(u) for u in us]
            g.step_vjp_pred_dev(uj, [p[j] for p in preds], lams, a, dts)
            g.step_vjp_dev(uj, lams, b, dts)
            torch.cuda.synchronize()
            for i in range(len(us)):
                worst_s = max(worst_s, rel_l2(host(trajs[i][j]), host(ref[i])))
                worst_v = max(worst_v, rel_l2(host(a[i]), host(b[i])))
        print(f"group record prec={prec}: states vs single steps {worst_s:.2e}, "
              f"VJP with recorded u* vs recomputed {worst_v:.2e}", flush=True)
        assert worst_s <= tol and worst_v <= tol


def test_group_overlapped_backward_equals_one_stream_loop(gpu_lib):
    """The batched checkpointed backward (twin group on a second stream, recorded u*) gives the
    gradients of a plain one-stream loop of single group steps and recomputing step adjoints."""
    import torch
    from dnsbox_amd import BoxGroup
    from dnsbox_amd.adjoint import group_step_map
    T, k = 30, 7
    us = fields(64, 70)
    gs = rand_like(us, 11)
    with BoxGroup([kw(n, nu, 64) for n, nu in zip(SIZES, NUS)]) as g:
        dts = [DT] * len(us)
        xs = [u.clone().requires_grad_(True) for u in us]
        got = torch.autograd.grad(group_step_map(g, xs, T, dts, k, record="checkpoint"), xs, gs)
        for b in g.boxes:
            b.rhs_int_fact(DT)
        ckpts, cur = [], [u.clone() for u in us]
        for t in range(0, T, k):
            ckpts.append([c.clone() for c in cur])
            g.step_dev(cur, min(k, T - t))
        lams, outs = [x.clone() for x in gs], [torch.empty_like(x) for x in gs]
        for seg in reversed(range(len(ckpts))):
            m = min(k, T - seg * k)
            states = [ckpts[seg]]
            for _ in range(m - 1):
                v = [c.clone() for c in states[-1]]
                g.step_dev(v, 1)
                states.append(v)
            for j in reversed(range(m)):
                g.step_vjp_dev(states[j], lams, outs, dts)
                lams, outs = outs, lams
        torch.cuda.synchronize()
    worst = max(rel_l2(host(a), host(b)) for a, b in zip(got, lams))
    print(f"group overlapped backward: worst rel_l2 vs one-stream loop {worst:.3e}", flush=True)
    assert worst <= 1e-14


@pytest.mark.parametrize("prec", [64, 32])
def test_group_vjp_cube_matches_full_passes(gpu_lib, prec):
    """Group step adjoints with set_vjp_cube (states made by pruned steps) agree with the full
    passes to rounding."""
    import torch
    from dnsbox_amd import BoxGroup
    tol = 1e-13 if prec == 64 else 1e-6
    us = fields(prec, 20)
    lams = rand_like(us, 21)
    with BoxGroup([kw(n, nu, prec) for n, nu in zip(SIZES, NUS)]) as g:
        for b in g.boxes:
            b.rhs_int_fact(DT)
        g.step_dev(us, 2)
        assert g.last_step_pruned()
        res = {}
        for cube in (False, True):
            g.set_vjp_cube(cube)
            outs = [torch.empty_like(u) for u in us]
            g.step_vjp_dev(us, lams, outs, [DT] * len(us))
            torch.cuda.synchronize()
            res[cube] = [host(o) for o in outs]
        g.set_vjp_cube(False)
    worst = max(rel_l2(a, b) for a, b in zip(res[True], res[False]))
    print(f"group vjp cube prec={prec}: worst rel {worst:.2e}", flush=True)
    assert worst <= tol


def test_group_vjp_cube_guard_reports_a_state_outside_the_cube(gpu_lib):
    """The group guarantee is checked on the device like the single box's: one box's state with
    energy outside its kept cube makes set_vjp_cube(False) raise; confined states pass."""
    import torch
    from dnsbox_amd import BoxGroup
    from dnsbox_amd._lib import DnsboxError
    us = fields(64, 30)
    lams = rand_like(us, 31)
    with BoxGroup([kw(n, nu, 64) for n, nu in zip(SIZES, NUS)]) as g:
        for b in g.boxes:
            b.rhs_int_fact(DT)
        g.step_dev(us, 1)
        assert g.last_step_pruned()
        outs = [torch.empty_like(u) for u in us]
        g.set_vjp_cube(True)
        g.step_vjp_dev(us, lams, outs, [DT] * len(us))
        g.set_vjp_cube(False)
        bad = list(us)
        bad[2] = rand_like([us[2]], 32)[0]   # box 2 (64^3): energy on every mode
        g.set_vjp_cube(True)
        g.step_vjp_dev(bad, lams, outs, [DT] * len(us))
        with pytest.raises(DnsboxError, match="outside"):
            g.set_vjp_cube(False)
    print("group vjp cube guard: confined states pass, a box outside its cube raises", flush=True)
