"""The step adjoint with recorded x-pass inputs (dnsbox_step_dev_record_mixed +
dnsbox_step_vjp_mixed_dev) against the step adjoint that transforms u and u* again
(dnsbox_step_vjp_pred_dev), on pruned and unpruned recordings, at sizes that run the plane
pass, the LDS-staged and the register x pass, and on a non-cubic grid; the recorded states
(stored by the corrector pass) against separately stepped ones; the segment entry
(dnsbox_step_vjp_mixed_seq_dev) against single step adjoints; the T-step-map gradient with and
without the recorded inputs; and the refusal of a pruned recording without the kept-cube
guarantee."""
import numpy as np
import pytest

from oracle_lib import Oracle, from_torch_layout, to_torch_layout

pytestmark = pytest.mark.gpu


def dev(f, prec=64):
    import torch
    t = torch.from_numpy(np.stack([to_torch_layout(a) for a in f])).cuda()
    return (t if prec == 64 else t.to(torch.complex64)).contiguous()


def host(t):
    import torch
    a = t.detach().cpu().to(torch.complex128).numpy()
    return [from_torch_layout(a[i]) for i in range(3)]


def rel_l2(a, b):
    num = sum(np.linalg.norm((x - y).ravel()) ** 2 for x, y in zip(a, b))
    return float(np.sqrt(num / sum(np.linalg.norm(y.ravel()) ** 2 for y in b)))


def rand_spec(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(3)]


CASES = [((32, 32, 32), 64), ((64, 64, 64), 64), ((48, 48, 48), 32), ((32, 48, 16), 64), ((256, 256, 256), 32),
         ((256, 256, 128), 32)]


@pytest.mark.parametrize("pruned", [True, False])
@pytest.mark.parametrize("grid,prec", CASES)
def test_mixed_step_vjp_equals_recomputed(gpu_lib, grid, prec, pruned):
    import torch
    from dnsbox_amd import Box
    nx, ny, nz = grid
    dt, nsteps = 0.004, 2
    o = Oracle(nx, ny=ny, nz=nz, nu=0.02, kolmo_amp=1.0, kolmo_k=3)
    with Box(n=nx, nx=nx, ny=ny, nz=nz, precision=prec, nu=0.02, kolmo_amp=1.0, kolmo_k=3, bandlim=0,
             adaptive_dt=0, dt=dt) as b:
        b.set_pruning(pruned)
        b.rhs_int_fact(dt)
        u0 = dev(o.init_solenoidal(3, 0, 3.0, 0.5), prec)
        lam = dev(rand_spec(o.spec_shape, 4), prec)
        traj = torch.empty((nsteps,) + tuple(u0.shape), dtype=u0.dtype, device=u0.device)
        preds = torch.full_like(traj, 7.0)   # stale values outside the kept cube must not survive
        traj.fill_(7.0)
        mixed = torch.empty((2 * nsteps,) + b.mixed_slot_shape, dtype=u0.dtype, device=u0.device)
        u = u0.clone()
        pr = b.step_dev_record_mixed(u, nsteps, traj, preds, mixed)
        assert pr == pruned
        ref = u0.clone()
        states = [u0]
        for j in range(nsteps):
            b.step_dev(ref, 1)
            assert rel_l2(host(traj[j]), host(ref)) <= (1e-15 if prec == 64 else 1e-7)
            states.append(ref.clone())
        assert torch.equal(u, traj[nsteps - 1])
        tol = 1e-13 if prec == 64 else 2e-6
        for cube in ([True, False] if not pruned else [True]):
            b.set_vjp_cube(cube)
            for j in range(nsteps):
                o1, o2 = torch.empty_like(lam), torch.empty_like(lam)
                b.step_vjp_pred_dev(states[j], preds[j], lam, o1, dt)
                b.step_vjp_mixed_dev(states[j], preds[j], lam, o2, dt, mixed[2 * j], mixed[2 * j + 1], pr)
                err = rel_l2(host(o2), host(o1))
                print(f"mixed step-VJP {grid} prec={prec} pruned={pruned} cube={cube} step {j}: rel_l2 {err:.3e}",
                      flush=True)
                assert err <= tol
            b.set_vjp_cube(False)


SEQ_CASES = [((32, 32, 32), 64, 3), ((64, 64, 64), 64, 4), ((48, 48, 48), 32, 3), ((32, 48, 16), 64, 4),
             ((256, 256, 256), 32, 2)]


@pytest.mark.parametrize("pruned", [True, False])
@pytest.mark.parametrize("grid,prec,nsteps", SEQ_CASES)
def test_mixed_seq_equals_single_steps(gpu_lib, grid, prec, nsteps, pruned):
    """dnsbox_step_vjp_mixed_seq_dev (a segment's step adjoints in one call; with the cube-confined
    passes each step's last z-forward also runs the step before's first pass) against
    dnsbox_step_vjp_mixed_dev step by step, odd and even step counts."""
    import torch
    from dnsbox_amd import Box
    nx, ny, nz = grid
    dt = 0.004
    o = Oracle(nx, ny=ny, nz=nz, nu=0.02, kolmo_amp=1.0, kolmo_k=3)
    with Box(n=nx, nx=nx, ny=ny, nz=nz, precision=prec, nu=0.02, kolmo_amp=1.0, kolmo_k=3, bandlim=0,
             adaptive_dt=0, dt=dt) as b:
        b.set_pruning(pruned)
        b.rhs_int_fact(dt)
        u0 = dev(o.init_solenoidal(5, 0, 3.0, 0.5), prec)
        lam0 = dev(rand_spec(o.spec_shape, 6), prec)
        traj = torch.empty((nsteps,) + tuple(u0.shape), dtype=u0.dtype, device=u0.device)
        preds = torch.empty_like(traj)
        mixed = torch.empty((2 * nsteps,) + b.mixed_slot_shape, dtype=u0.dtype, device=u0.device)
        pr = b.step_dev_record_mixed(u0.clone(), nsteps, traj, preds, mixed)
        states = [u0] + [traj[j] for j in range(nsteps - 1)]
        tol = 1e-13 if prec == 64 else 2e-6
        for cube in ([True, False] if not pruned else [True]):
            b.set_vjp_cube(cube)
            lam, out = lam0.clone(), torch.empty_like(lam0)
            for j in reversed(range(nsteps)):
                b.step_vjp_mixed_dev(states[j], preds[j], lam, out, dt, mixed[2 * j], mixed[2 * j + 1], pr)
                lam, out = out, lam
            l2, o2 = lam0.clone(), torch.full_like(lam0, 3.0)
            b.step_vjp_mixed_seq_dev(u0, traj, preds, l2, o2, dt, mixed, nsteps, pr)
            err = rel_l2(host(o2), host(lam))
            print(f"mixed seq {grid} prec={prec} n={nsteps} pruned={pruned} cube={cube}: rel_l2 {err:.3e}",
                  flush=True)
            assert err <= tol
            b.set_vjp_cube(False)


def test_pruned_recording_needs_the_cube_guarantee(gpu_lib):
    import torch
    from dnsbox_amd import Box
    from dnsbox_amd._lib import DnsboxError
    n, dt = 32, 0.004
    o = Oracle(n, nu=0.02)
    with Box(n=n, precision=64, nu=0.02, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.rhs_int_fact(dt)
        u = dev(o.init_solenoidal(3, 0, 3.0, 0.5))
        traj = torch.empty((1,) + tuple(u.shape), dtype=u.dtype, device=u.device)
        preds = torch.empty_like(traj)
        mixed = torch.empty((2,) + b.mixed_slot_shape, dtype=u.dtype, device=u.device)
        u0 = u.clone()
        assert b.step_dev_record_mixed(u, 1, traj, preds, mixed)
        out = torch.empty_like(u)
        with pytest.raises(DnsboxError, match="set_vjp_cube"):
            b.step_vjp_mixed_dev(u0, preds[0], u0, out, dt, mixed[0], mixed[1], True)


@pytest.mark.parametrize("record,k", [("all", 10), ("checkpoint", 7)])
def test_step_map_gradient_with_recorded_inputs(gpu_lib, record, k):
    """The T-step-map gradient with the x-pass inputs recorded equals the one that transforms the
    states again (RECORD_MIXED off), whole trajectory recorded or checkpointed."""
    import torch
    from dnsbox_amd import Box, adjoint
    n, dt, T = 32, 0.005, 30
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    with Box(n=n, precision=64, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        u0 = dev(o.init_solenoidal(8, 0, 4.0, 0.5))
        g = dev(rand_spec(o.spec_shape, 9))
        grads = {}
        saved = adjoint.RECORD_MIXED
        try:
            for mixed in (True, False):
                adjoint.RECORD_MIXED = mixed
                x = u0.clone().requires_grad_(True)
                (grads[mixed],) = torch.autograd.grad(adjoint.step_map(b, x, T, dt, k, record=record), x, g)
        finally:
            adjoint.RECORD_MIXED = saved
        torch.cuda.synchronize()
    err = rel_l2(host(grads[True]), host(grads[False]))
    print(f"T-step gradient, recorded inputs vs recomputed ({record}): rel_l2 {err:.3e}", flush=True)
    assert err <= 1e-13


GROUP = [(32, 0.02), (48, 0.03), (64, 0.015), (16, 0.05), (32, 0.04)]


def group_kw(n, nu, prec):
    return dict(n=n, precision=prec, nu=nu, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.005)


@pytest.mark.parametrize("prec", [32, 64])
def test_group_mixed_step_vjp_equals_recomputed(gpu_lib, prec):
    """The group form: dnsbox_group_step_dev_record_mixed stores the same states as separate group
    steps, and dnsbox_group_step_vjp_mixed_dev equals dnsbox_group_step_vjp_pred_dev per box."""
    import torch
    from dnsbox_amd import BoxGroup
    dt, nsteps = 0.005, 2
    tol = 1e-13 if prec == 64 else 2e-6
    with BoxGroup([group_kw(n, nu, prec) for n, nu in GROUP]) as g:
        for b in g.boxes:
            b.rhs_int_fact(dt)
        us = [dev(Oracle(n, nu=nu).init_solenoidal(20 + i, 0, 3.0, 0.5), prec) for i, (n, nu) in enumerate(GROUP)]
        lams = [dev(rand_spec(Oracle(n).spec_shape, 30 + i), prec) for i, (n, _) in enumerate(GROUP)]
        trajs = [torch.full((nsteps,) + tuple(u.shape), 7.0, dtype=u.dtype, device=u.device) for u in us]
        preds = [torch.full_like(t, 7.0) for t in trajs]
        mixeds = [torch.empty((2 * nsteps,) + b.mixed_slot_shape, dtype=u.dtype, device=u.device)
                  for b, u in zip(g.boxes, us)]
        st = [u.clone() for u in us]
        assert g.step_dev_record_mixed(st, nsteps, trajs, preds, mixeds)
        ref = [u.clone() for u in us]
        states = [[u.clone() for u in us]]
        for j in range(nsteps):
            g.step_dev(ref, 1)
            states.append([r.clone() for r in ref])
            for i in range(len(us)):
                assert rel_l2(host(trajs[i][j]), host(ref[i])) <= (1e-15 if prec == 64 else 1e-7)
        g.set_vjp_cube(True)
        for j in range(nsteps):
            o1 = [torch.empty_like(u) for u in us]
            o2 = [torch.empty_like(u) for u in us]
            g.step_vjp_pred_dev(states[j], [p[j] for p in preds], lams, o1, [dt] * len(us))
            g.step_vjp_mixed_dev(states[j], [p[j] for p in preds], lams, o2, [dt] * len(us),
                                 [mx[2 * j:2 * j + 2] for mx in mixeds], True)
            for i in range(len(us)):
                err = rel_l2(host(o2[i]), host(o1[i]))
                print(f"group mixed step-VJP prec={prec} step {j} box {i} (n={GROUP[i][0]}): rel_l2 {err:.3e}",
                      flush=True)
                assert err <= tol
        g.set_vjp_cube(False)


def test_group_gradients_with_recorded_inputs(gpu_lib):
    """group_residual_grads with the recorded inputs equals the recomputing backward (FP32 group,
    whole trajectory recorded and checkpointed)."""
    import torch
    from dnsbox_amd import BoxGroup, adjoint
    dt, T = 0.005, 24
    with BoxGroup([group_kw(n, nu, 32) for n, nu in GROUP]) as g:
        us = [dev(Oracle(n, nu=nu).init_solenoidal(50 + i, 0, 3.0, 0.5), 32) for i, (n, nu) in enumerate(GROUP)]
        saved = adjoint.RECORD_MIXED
        for record, k in (("all", 8), ("checkpoint", 7)):
            grads = {}
            try:
                for mixed in (True, False):
                    adjoint.RECORD_MIXED = mixed
                    _, grads[mixed] = adjoint.group_residual_grads(g, us, T, [dt] * len(us), k, record)
            finally:
                adjoint.RECORD_MIXED = saved
            for i in range(len(us)):
                err = rel_l2(host(grads[True][i]), host(grads[False][i]))
                print(f"group gradient ({record}) box {i}: rel_l2 {err:.3e}", flush=True)
                assert err <= 1e-5
