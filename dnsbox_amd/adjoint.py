"""Differentiable T-step map and adjoint descent on periodic-orbit residuals.

``step_map(box, u0, T, dt, k)`` returns u_T = Phi_T(u0) for T fixed-dt IF-Heun steps
(dnsboxheun.f90:84-126) as a ``torch.autograd.Function``.  When a gradient is needed the
forward records the trajectory (every state and every step's predicted state) in device
memory if its 2 (T - 1) states fit in a quarter of the free device memory (record="auto"),
and the backward runs the hand-written step-adjoint kernels (``dnsbox_step_vjp_dev``)
backwards through it with no recomputation; otherwise (or with record="checkpoint") the
forward keeps a checkpoint every k steps and the backward recomputes each segment from its
checkpoint.  Gradients follow torch's convention for complex tensors (dL/dRe + i dL/dIm),
i.e. the Euclidean adjoint in the stored coordinates.

``residual`` is J(u) = 1/2 <|Phi_T(u) - u|^2> (the mean-square physical residual, in
FP64), ``adjoint_descent`` minimises it with an FP64 backtracking line search along
the projected L2-metric gradient.  This is the T-step-map counterpart of the
reference's equilibrium adjoint descent (dnsboxAdj.f90:76-126 with rhsAdj.f90).
"""
from __future__ import annotations

import numpy as np
import torch


def _alphas(box):
    """alpha_x, alpha_y, alpha_z of a box as the library holds them (single precision, 0 = 1)."""
    p = box.params
    return tuple(float(np.float32(a)) if a > 0 else 1.0 for a in (p.alpha_x, p.alpha_y, p.alpha_z))


def _k_tables(box, device):
    """The scaled table wavenumbers (Nyquist 0) of the internal layout (nz, ny, nx/2+1)."""
    def table(n, a):
        k = torch.arange(n, device=device, dtype=torch.float64)
        return torch.where(k < n // 2, k, torch.where(k == n // 2, torch.zeros_like(k), k - n)) * a
    ax, ay, az = _alphas(box)
    kz = table(box.nz, az).view(box.nz, 1, 1)
    ky = table(box.ny, ay).view(1, box.ny, 1)
    kx = table(box.nx, ax)[: box.nx // 2 + 1].view(1, 1, box.nx // 2 + 1)
    return kx, ky, kz


_WEIGHTS = {}


def hermitian_weight(n, device):
    """1 on the kx = 0 and kx = NX/2 planes, 2 elsewhere (c2r counts those twice); n = NX.  One
    tensor per (n, device), made once (callers only read it): three fills per residual otherwise
    stand between the forward and the step adjoints."""
    key = (n, str(torch.device(device)))
    w = _WEIGHTS.get(key)
    if w is None:
        w = torch.full((1, 1, n // 2 + 1), 2.0, dtype=torch.float64, device=device)
        w[..., 0] = 1.0
        w[..., n // 2] = 1.0
        _WEIGHTS[key] = w
    return w


def _require_fixed_dt(boxes):
    """The step adjoints differentiate fixed-dt steps: a box whose CFL controller would pick the
    step size itself (adaptive_dt) cannot be stepped by the T-step map."""
    for b in boxes:
        if b.params.adaptive_dt:
            raise ValueError("the T-step map needs boxes created with adaptive_dt=0 (fixed dt): with the "
                             "CFL controller the forward steps would not be the steps the adjoint differentiates")


class _StepMap(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, box, T, dt, k, record):
        _require_fixed_dt([box])
        box.rhs_int_fact(dt)
        u = u0.detach().clone().contiguous()
        if ctx.needs_input_grad[0]:
            k = _segment_length(T, k, u.numel() * u.element_size(), u.device, record)
        ckpts = []
        cube = True
        ctx.last = None
        starts = list(range(0, T, k))
        for t in starts:   # k steps per call: the step pipeline runs through a segment
            ckpts.append(u.clone())
            m = min(k, T - t)
            if t == starts[-1] and m > 1 and ctx.needs_input_grad[0]:
                # the last segment, which the backward starts from, is recorded now rather than
                # recomputed then: its first m - 1 states and predicted states (and, memory
                # permitting, the x-pass inputs of their right-hand sides), then its last step
                shape = (max(k - 1, 1),) + tuple(u.shape)
                traj = torch.empty(shape, dtype=u.dtype, device=u.device)
                preds = torch.empty(shape, dtype=u.dtype, device=u.device)
                fits = _mixed_fits(box, max(k - 1, 1), u.device)
                mixed = _mixed_buffer(box, max(k - 1, 1), u.device) if fits else None
                if mixed is not None:
                    ctx.last = (traj, preds, mixed, box.step_dev_record_mixed(u, m - 1, traj, preds, mixed))
                else:
                    box.step_dev_record(u, m - 1, traj, preds)
                    ctx.last = (traj, preds, None, False)
                cube = cube and box.last_step_pruned()
                box.step_dev(u, 1)
            else:
                box.step_dev(u, m)
            cube = cube and box.last_step_pruned()
        ctx.box, ctx.T, ctx.dt, ctx.k = box, T, dt, k
        ctx.ckpts = ckpts
        ctx.cube = cube   # every state of the trajectory lies in the kept cube
        return u

    @staticmethod
    def backward(ctx, g):
        box, T, dt, k = ctx.box, ctx.T, ctx.dt, ctx.k
        box.rhs_int_fact(dt)
        lam = g.detach().clone().contiguous()
        out = torch.empty_like(lam)
        nseg = len(ctx.ckpts)
        shape = (max(k - 1, 1),) + tuple(lam.shape)
        # Segment s's states u_start .. u_{stop-1} and the predicted states of its steps but the
        # last (and, memory permitting, the x-pass inputs of their right-hand sides) are
        # recomputed by a twin box (same parameters, own buffers) on a second stream into one of
        # two buffer sets, while this box runs the step adjoints of segment s + 1 from the other:
        # at small N neither fills the GPU, so the two overlap.
        nbuf = min(2, nseg) - (1 if ctx.last is not None else 0)
        mix = nseg > 1 and _mixed_fits(box, max(k - 1, 1) * nbuf, lam.device)
        bufs = [(torch.empty(shape, dtype=lam.dtype, device=lam.device),
                 torch.empty(shape, dtype=lam.dtype, device=lam.device),
                 _mixed_buffer(box, max(k - 1, 1), lam.device) if mix else None)
                for _ in range(nbuf)]
        rec_pruned = [False] * nseg       # segment s's recording ran pruned
        if ctx.last is not None:   # the last segment was recorded by the forward
            bufs.insert((nseg - 1) % 2, ctx.last[:3])
            rec_pruned[nseg - 1] = ctx.last[3]
        main = torch.cuda.current_stream(lam.device)
        if nseg > 1 or ctx.last is None:  # segments to recompute
            rec = box.twin()
            rec.rhs_int_fact(dt)
            side = _side_stream(box, lam.device)
            side.wait_stream(main)        # the checkpoints and buffers are ready
        recorded = [None] * nseg          # segment s recomputed into bufs[s % 2]
        adjointed = [None] * nseg         # segment s's step adjoints done with bufs[s % 2]

        def span(seg):
            return seg * k, min(seg * k + k, T)

        def recompute(seg):
            start, stop = span(seg)
            with torch.cuda.stream(side):
                if seg + 2 < nseg:
                    side.wait_event(adjointed[seg + 2])   # its buffers were segment seg + 2's
                if stop - start > 1:
                    traj, preds, mixed = bufs[seg % 2]
                    if mixed is not None:
                        rec_pruned[seg] = rec.step_dev_record_mixed(ctx.ckpts[seg].clone(), stop - start - 1, traj,
                                                                    preds, mixed)
                    else:
                        rec.step_dev_record(ctx.ckpts[seg].clone(), stop - start - 1, traj, preds)
                recorded[seg] = torch.cuda.Event()
                recorded[seg].record(side)

        if ctx.last is not None:
            recorded[nseg - 1] = torch.cuda.Event()
            recorded[nseg - 1].record(main)
        else:
            recompute(nseg - 1)
        box.set_vjp_cube(ctx.cube)   # states made by pruned steps: the adjoints read the cube only
        try:
            for seg in reversed(range(nseg)):
                main.wait_event(recorded[seg])
                start, stop = span(seg)
                traj, preds, mixed = bufs[seg % 2]
                u0 = ctx.ckpts[seg]   # step j of the segment starts at u0 (j = 0) or traj[j - 1]
                # a pruned recording holds the kept kx only: its inputs serve the cube-confined adjoints
                use_mixed = mixed is not None and (ctx.cube or not rec_pruned[seg])
                # the segment's last step recomputes its own u*; the steps before it read the
                # recorded u* (and x-pass inputs, in one call for the whole run of steps)
                last = stop - start - 1
                box.step_vjp_dev(traj[last - 1] if last > 0 else u0, lam, out, dt)
                lam, out = out, lam
                if last > 0 and use_mixed:
                    box.step_vjp_mixed_seq_dev(u0, traj, preds, lam, out, dt, mixed, last, rec_pruned[seg])
                    lam, out = out, lam
                else:
                    for j in reversed(range(last)):
                        box.step_vjp_pred_dev(traj[j - 1] if j > 0 else u0, preds[j], lam, out, dt)
                        lam, out = out, lam
                adjointed[seg] = torch.cuda.Event()
                adjointed[seg].record(main)
                if seg > 0:
                    recompute(seg - 1)
        finally:
            box.set_vjp_cube(False)
        return lam, None, None, None, None, None


RECORD_FRACTION = 0.25   # share of the free device memory a recorded trajectory may take
RECORD_MIXED = True      # record the x-pass inputs of recorded steps for their adjoints


def _mixed_fits(box, nsteps, device):
    """Whether record slots for the x-pass inputs of nsteps steps (two 9-field slots per step) are
    wanted (RECORD_MIXED) and fit in RECORD_FRACTION of the free device memory."""
    if not RECORD_MIXED or nsteps < 1:
        return False
    need = 2 * nsteps * int(np.prod(box.mixed_slot_shape)) * (16 if box.precision == 64 else 8)
    free, _ = torch.cuda.mem_get_info(device)
    return need <= RECORD_FRACTION * free


def _mixed_buffer(box, nsteps, device):
    """Record slots for nsteps steps (step_dev_record_mixed)."""
    dtype = torch.complex128 if box.precision == 64 else torch.complex64
    return torch.empty((2 * nsteps,) + tuple(box.mixed_slot_shape), dtype=dtype, device=device)


def _group_mixed(group, nsteps, device, allocate=True, check=True):
    """Record slots of nsteps steps for every box of a group (a list of tensors) when RECORD_MIXED
    is set and they fit in RECORD_FRACTION of the free device memory (check), else None; with
    allocate=False only whether they fit."""
    if check:
        if not RECORD_MIXED or nsteps < 1:
            return None if allocate else False
        need = sum(2 * nsteps * int(np.prod(b.mixed_slot_shape)) * (16 if b.precision == 64 else 8)
                   for b in group.boxes)
        free, _ = torch.cuda.mem_get_info(device)
        if need > RECORD_FRACTION * free:
            return None if allocate else False
    if not allocate:
        return True
    return [_mixed_buffer(b, nsteps, device) for b in group.boxes]


def _segment_length(T, k, state_bytes, device, record):
    """Steps per segment of the forward: T (the whole trajectory recorded, no recomputation in
    the backward) for record="all", or for "auto" when the 2 (T - 1) recorded states of
    state_bytes each fit in RECORD_FRACTION of the free device memory; else the checkpoint
    interval k."""
    if record not in ("auto", "all", "checkpoint"):
        raise ValueError(f"record must be 'auto', 'all' or 'checkpoint', not {record!r}")
    if k < 1:
        raise ValueError("the checkpoint interval k must be >= 1")
    if record == "checkpoint" or T <= 1:
        return k
    if record == "all":
        return T
    free, _ = torch.cuda.mem_get_info(device)
    return T if 2 * (T - 1) * state_bytes <= RECORD_FRACTION * free else k


def _side_stream(owner, device):
    """The second stream of owner's checkpointed backwards, created once: creating a stream costs
    milliseconds of host time, during which the device would idle between the forward and the
    step adjoints."""
    s = getattr(owner, "_adj_side_stream", None)
    if s is None or s.device != device:
        s = torch.cuda.Stream(device=device)
        owner._adj_side_stream = s
    return s


def _copy_all(dst, src):
    torch._foreach_copy_(dst, src)


class _GroupStepMap(torch.autograd.Function):
    """Phi_T of every box of a BoxGroup at once (one launch per pass for all boxes)."""

    @staticmethod
    def forward(ctx, group, T, dts, k, record, *u0s):
        _require_fixed_dt(group.boxes)
        for b, dt in zip(group.boxes, dts):
            b.rhs_int_fact(dt)
        us = [u.detach().clone().contiguous() for u in u0s]
        if any(ctx.needs_input_grad[5:]):
            k = _segment_length(T, k, sum(u.numel() * u.element_size() for u in us), us[0].device, record)
        ckpts = []
        cube = True
        ctx.last = None
        starts = list(range(0, T, k))
        for t in starts:   # k steps per call: the step pipeline runs through a segment
            c = [torch.empty_like(u) for u in us]
            _copy_all(c, us)
            ckpts.append(c)
            m = min(k, T - t)
            if t == starts[-1] and m > 1 and any(ctx.needs_input_grad[5:]):
                # the last segment is recorded now (as in _StepMap.forward)
                w = max(k - 1, 1)
                pairs = [(torch.empty((w,) + tuple(u.shape), dtype=u.dtype, device=u.device),
                          torch.empty((w,) + tuple(u.shape), dtype=u.dtype, device=u.device)) for u in us]
                mixeds = _group_mixed(group, w, us[0].device)
                if mixeds is not None:
                    pr = group.step_dev_record_mixed(us, m - 1, [a for a, _ in pairs], [b for _, b in pairs], mixeds)
                else:
                    group.step_dev_record(us, m - 1, [a for a, _ in pairs], [b for _, b in pairs])
                    pr = False
                ctx.last = (pairs, mixeds, pr)
                cube = cube and group.last_step_pruned()
                group.step_dev(us, 1)
            else:
                group.step_dev(us, m)
            cube = cube and group.last_step_pruned()
        ctx.group, ctx.T, ctx.dts, ctx.k, ctx.ckpts = group, T, list(dts), k, ckpts
        ctx.cube = cube
        return tuple(us)

    @staticmethod
    def backward(ctx, *gs):
        group, T, dts, k = ctx.group, ctx.T, ctx.dts, ctx.k
        for b, dt in zip(group.boxes, dts):
            b.rhs_int_fact(dt)
        lams = [torch.zeros_like(c) if g is None else g.detach().clone().contiguous()
                for g, c in zip(gs, ctx.ckpts[0])]
        outs = [torch.empty_like(l) for l in lams]
        nseg = len(ctx.ckpts)
        # as in _StepMap.backward: a twin group recomputes segment s on a second stream (states
        # and predicted states recorded in one call) into one of two buffer sets while this group
        # runs the adjoints of segment s + 1 from the other
        m = max(k - 1, 1)
        nbuf = min(2, nseg) - (1 if ctx.last is not None else 0)
        mix = nseg > 1 and _group_mixed(group, m * nbuf, lams[0].device, allocate=False)
        # a buffer set: per box (states, predicted states), and the record slots or None
        bufs = [([(torch.empty((m,) + tuple(l.shape), dtype=l.dtype, device=l.device),
                   torch.empty((m,) + tuple(l.shape), dtype=l.dtype, device=l.device)) for l in lams],
                 _group_mixed(group, m, lams[0].device, check=False) if mix else None)
                for _ in range(nbuf)]
        rec_pruned = [False] * nseg
        if ctx.last is not None:   # the last segment was recorded by the forward
            bufs.insert((nseg - 1) % 2, ctx.last[:2])
            rec_pruned[nseg - 1] = ctx.last[2]
        main = torch.cuda.current_stream(lams[0].device)
        if nseg > 1 or ctx.last is None:  # segments to recompute
            rec = group.twin()
            for b, dt in zip(rec.boxes, dts):
                b.rhs_int_fact(dt)
            side = _side_stream(group, lams[0].device)
            side.wait_stream(main)
        recorded, adjointed = [None] * nseg, [None] * nseg

        def span(seg):
            return seg * k, min(seg * k + k, T)

        def recompute(seg):
            start, stop = span(seg)
            with torch.cuda.stream(side):
                if seg + 2 < nseg:
                    side.wait_event(adjointed[seg + 2])
                if stop - start > 1:
                    us = [c.clone() for c in ctx.ckpts[seg]]
                    pairs, mixeds = bufs[seg % 2]
                    if mixeds is not None:
                        rec_pruned[seg] = rec.step_dev_record_mixed(us, stop - start - 1, [t for t, _ in pairs],
                                                                    [p for _, p in pairs], mixeds)
                    else:
                        rec.step_dev_record(us, stop - start - 1, [t for t, _ in pairs], [p for _, p in pairs])
                recorded[seg] = torch.cuda.Event()
                recorded[seg].record(side)

        if ctx.last is not None:
            recorded[nseg - 1] = torch.cuda.Event()
            recorded[nseg - 1].record(main)
        else:
            recompute(nseg - 1)
        group.set_vjp_cube(ctx.cube)
        try:
            for seg in reversed(range(nseg)):
                main.wait_event(recorded[seg])
                start, stop = span(seg)
                buf, mixeds = bufs[seg % 2]
                use_mixed = mixeds is not None and (ctx.cube or not rec_pruned[seg])
                for j in reversed(range(stop - start)):
                    states = ctx.ckpts[seg] if j == 0 else [t[j - 1] for t, _ in buf]
                    if j < stop - start - 1 and use_mixed:
                        group.step_vjp_mixed_dev(states, [p[j] for _, p in buf], lams, outs, dts,
                                                 [mx[2 * j:2 * j + 2] for mx in mixeds], rec_pruned[seg])
                    elif j < stop - start - 1:
                        group.step_vjp_pred_dev(states, [p[j] for _, p in buf], lams, outs, dts)
                    else:
                        group.step_vjp_dev(states, lams, outs, dts)
                    lams, outs = outs, lams
                adjointed[seg] = torch.cuda.Event()
                adjointed[seg].record(main)
                if seg > 0:
                    recompute(seg - 1)
        finally:
            group.set_vjp_cube(False)
        return (None, None, None, None, None, *lams)


def group_step_map(group, u0s, T: int, dts, k: int = 20, record: str = "auto"):
    """[Phi_T(u0_i)] for every box i of a BoxGroup, differentiable; the trajectory is recorded
    or checkpointed every k as in step_map."""
    return _GroupStepMap.apply(group, T, list(dts), k, record, *u0s)


def group_residuals(group, u0s, T, dts, k=20, record="auto"):
    """[J_i] = [1/2 <|Phi_T(u0_i) - u0_i|^2>] for every box (FP64 scalars, one tensor)."""
    uTs = group_step_map(group, u0s, T, dts, k, record)
    return torch.stack([0.5 * mean_square(uT - u0, b)
                        for uT, u0, b in zip(uTs, u0s, group.boxes)])


def group_residual_grads(group, u0s, T, dts, k=20, record="auto"):
    """(J, [dJ_i/du0_i]): the residuals and their Euclidean gradients, all boxes at once."""
    xs = [u.detach().clone().requires_grad_(True) for u in u0s]
    J = group_residuals(group, xs, T, dts, k, record)
    gs = torch.autograd.grad(J.sum(), xs)
    return J.detach(), list(gs)


def group_adjoint_descent(group, u0s, T, dts, k=20, iters=10, step0=1.0, c1=1e-4, shrink=0.5,
                          max_backtrack=20, callback=None, record="auto"):
    """adjoint_descent for every box of a group at once: one batched gradient per iteration,
    an FP64 Armijo backtracking line search per box (each box keeps its own step length;
    trial residuals of all boxes are evaluated in one batched forward).  Returns (us, hist)
    with hist[it] the list of residuals of the boxes."""
    ws = [hermitian_weight(b.nx, u.device) for b, u in zip(group.boxes, u0s)]
    us = [u.detach().clone() for u in u0s]
    nb = len(us)
    alpha = [step0] * nb
    hist = []
    for it in range(iters):
        J, gs = group_residual_grads(group, us, T, dts, k, record)
        J0 = J.tolist()
        hist.append(J0)
        if callback:
            callback(it, J0)
        ds, slopes = [], []
        for b, g, w in zip(group.boxes, gs, ws):
            gm = g.to(torch.complex128) / w * float(b.nx * b.ny * b.nz) ** 2
            d = -project_masked(gm, b).to(g.dtype)
            ds.append(d)
            slopes.append(float((g.to(torch.complex128).conj() * d.to(torch.complex128)).real.sum()))
        pending = list(range(nb))
        trial = list(us)
        for _ in range(max_backtrack):
            if not pending:
                break
            for i in pending:
                trial[i] = us[i] + alpha[i] * ds[i]
            with torch.no_grad():
                Jt = group_residuals(group, trial, T, dts, k, record).tolist()
            still = []
            for i in pending:
                if Jt[i] <= J0[i] + c1 * alpha[i] * slopes[i]:
                    us[i] = trial[i]
                    alpha[i] *= 2.0
                else:
                    alpha[i] *= shrink
                    trial[i] = us[i]
                    still.append(i)
            pending = still
    with torch.no_grad():
        hist.append(group_residuals(group, us, T, dts, k, record).tolist())
    return us, hist


def step_map(box, u0, T: int, dt: float, k: int = 20, record: str = "auto"):
    """Phi_T(u0): T fixed-dt IF-Heun steps, differentiable.  record="auto" records the whole
    trajectory for the backward when it fits in RECORD_FRACTION of the free device memory and
    checkpoints every k steps otherwise; "all" always records, "checkpoint" always checkpoints."""
    return _StepMap.apply(u0, box, T, dt, k, record)


def mean_square(r, box):
    """<|r_phys|^2> of a spectral field (3, nz, ny, nx/2+1) of a box, in FP64 (Parseval, Hermitian
    weights)."""
    w = hermitian_weight(box.nx, r.device)
    rr = r.to(torch.complex128)
    return (w * (rr.real ** 2 + rr.imag ** 2)).sum() / float(box.nx * box.ny * box.nz) ** 2


def residual(box, u0, T, dt, k=20, record="auto"):
    """J(u0) = 1/2 <|Phi_T(u0) - u0|^2>."""
    uT = step_map(box, u0, T, dt, k, record)
    return 0.5 * mean_square(uT - u0, box)


def project_masked(g, box, parity=True):
    """Leray projection and cubic 2/3 mask (state_project + state_dealias, state.f90:120-209) of a
    field of a box: |k_d| >= (N_d/2)(2 alpha_d/3) on the scaled wavenumbers, as the library's mask."""
    kx, ky, kz = _k_tables(box, g.device)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    ax, ay, az = _alphas(box)
    m = ((kx.abs() >= (box.nx / 2.0) * (2.0 * ax / 3.0)) | (ky.abs() >= (box.ny / 2.0) * (2.0 * ay / 3.0))
         | (kz.abs() >= (box.nz / 2.0) * (2.0 * az / 3.0)) | (k2 == 0))
    if not parity:
        def nyq(n):
            return torch.arange(n, device=g.device) == n // 2
        m = (m | nyq(box.nz).view(box.nz, 1, 1) | nyq(box.ny).view(1, box.ny, 1)
             | nyq(box.nx)[: box.nx // 2 + 1].view(1, 1, box.nx // 2 + 1))
    k2s = torch.where(k2 == 0, torch.ones_like(k2), k2)
    gg = g.to(torch.complex128)
    kd = (kx * gg[0] + ky * gg[1] + kz * gg[2]) / k2s
    out = torch.stack([gg[0] - kx * kd, gg[1] - ky * kd, gg[2] - kz * kd])
    return torch.where(m.unsqueeze(0), torch.zeros_like(out), out).to(g.dtype)


def adjoint_descent(box, u0, T, dt, k=20, iters=50, step0=1.0, c1=1e-4, shrink=0.5,
                    max_backtrack=20, callback=None, record="auto"):
    """Minimise J(u) = 1/2 <|Phi_T(u) - u|^2> from u0 by gradient descent in the L2 metric
    (W^-1 times the Euclidean gradient, projected onto retained solenoidal fields) with an
    Armijo backtracking line search in FP64.  Returns (u, history of J)."""
    w = hermitian_weight(box.nx, u0.device)
    u = u0.detach().clone()
    hist = []
    alpha = step0
    for it in range(iters):
        x = u.clone().requires_grad_(True)
        J = residual(box, x, T, dt, k, record)
        (g,) = torch.autograd.grad(J, x)
        # J = 1/2 sum W |r|^2 / M^2 (M = nx ny nz points) in the metric <a, b>_M = sum W Re(conj a b)
        # / M^2, whose gradient is M^2 W^-1 times the Euclidean one; descend along its projection
        gm = g.to(torch.complex128) / w * float(box.nx * box.ny * box.nz) ** 2
        d = -project_masked(gm, box).to(u.dtype)
        slope = float((g.to(torch.complex128).conj() * d.to(torch.complex128)).real.sum())
        J0 = float(J.detach())
        hist.append(J0)
        if callback:
            callback(it, J0)
        accepted = False
        for _ in range(max_backtrack):
            trial = u + alpha * d
            with torch.no_grad():
                Jt = float(residual(box, trial, T, dt, k, record))
            if Jt <= J0 + c1 * alpha * slope:
                u = trial
                accepted = True
                alpha *= 2.0
                break
            alpha *= shrink
        if not accepted:
            break
    with torch.no_grad():
        hist.append(float(residual(box, u, T, dt, k, record)))
    return u, hist
