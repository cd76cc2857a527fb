#!/usr/bin/env python3
"""Benchmark of the MI355X pseudospectral Navier-Stokes step (one GPU).

Workload of the headline number: 256^3 box, FP32 FFTs and nonlinear term, FP64 state
update, integrating factor, CFL control and diagnostics; nu = 0.005, Kolmogorov forcing
f = (sin 4y, 0, 0), IF-Heun predictor-corrector with the CFL controller at Courant 0.5,
seeded solenoidal initial field E(k) ~ k^4 exp(-2 (k/4)^2) normalised to E = 0.5.  A step
is one IF-Heun step (two RHS evaluations) of the whole box, state resident in HBM.

Printed as one JSON line:
  value                  steps/s of that FP32 form
  fp32_pure_*            the all-FP32 form (FP32 state too): rate and error
  fp64_steps_per_s       the same workload in FP64 (speedup_vs_fp64: value over it)
  fp64_unpruned_*        the FP64 step with whole-field transforms, the reference's transform
                         structure (speedup_vs_fp64_unpruned: value over it)
  *_err_vs_fp64          spectrum (per shell and L2) and energy error vs the FP64 run after
                         --err-steps fixed-dt steps from the same field
  unpruned_steps_per_s   the headline form with the transforms unpruned (the step prunes its
                         passes to the 2/3-rule cube; results agree to rounding)
  roofline               the dominant kernel of the step: algorithmic bytes (pruned) / HIP-event
                         launch time vs the 8 TB/s HBM peak; traffic = measured HBM bytes
                         per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)
                         when a measurement of this kernel is on file
  large_box              the same precision form at 512^3 (nu = 0.002): FP32-form and FP64
                         steps/s and the FP32 form's error vs FP64 after 1000 fixed-dt steps
  adjoint                64^3 FP64 checkpointed T-step gradients per second (T = 200,
                         dt = 0.005, checkpoint every 20), one finite-difference check, and
                         50 adjoint-descent iterations with the FP64 line search (J0 -> J50)
  batched                48 boxes (16 x 32^3, 16 x 48^3, 16 x 64^3, FP32, own nu and k0):
                         box-steps/s forward (hipGraph step) and forward + adjoint
                         (checkpointed gradient of the T = 100 map)
  cpu_baseline           the FP64 CPU oracle on the host cores (a reported baseline)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-steps/s at 256³ FP32 vs FP64 (err stated), % HBM roofline, "
          "adjoint grads/s")
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_256cubed.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nu", type=float, default=None, help="viscosity (default 0.005 at n <= 256, else 0.002)")
    ap.add_argument("--err-steps", type=int, default=1000,
                    help="fixed-dt steps of the FP32-vs-FP64 error runs (0 = skip)")
    ap.add_argument("--fp64-steps", type=int, default=300, help="timed FP64 steps")
    ap.add_argument("--pure-steps", type=int, default=1000, help="timed all-FP32 steps (0 = skip)")
    ap.add_argument("--unpruned-steps", type=int, default=300,
                    help="timed steps of the headline form with unpruned transforms (0 = skip)")
    ap.add_argument("--adjoint", type=int, default=1, help="time adjoint gradients (1/0)")
    ap.add_argument("--descent-iters", type=int, default=50,
                    help="adjoint-descent iterations run after the gradient timing (0 = skip)")
    ap.add_argument("--batched", type=int, default=1, help="time the 48-box group (1/0)")
    ap.add_argument("--grids", type=int, default=1, help="time the non-cubic grids (1/0)")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle (1/0)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--reps", type=int, default=20, help="launches per pass for kernel timing")
    ap.add_argument("--large-n", type=int, default=512,
                    help="the larger precision-form box (nu = 0.002), reported beside the headline (0 = skip)")
    ap.add_argument("--large-steps", type=int, default=100, help="timed steps of the larger box")
    ap.add_argument("--large-err-steps", type=int, default=1000,
                    help="fixed-dt steps of the larger box's FP32-vs-FP64 error runs (0 = skip)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


NU = 0.005


def make_box(n, precision, adaptive=True, dt=0.002, state_precision=0):
    from dnsbox_amd import Box
    return Box(n=n, precision=precision, nu=NU, kolmo_amp=1.0, kolmo_k=4, bandlim=0,
               adaptive_dt=int(adaptive), cfl_min=0.5, cfl_max=0.5, dt_max=0.01, dt=dt,
               state_precision=state_precision)


def timed_steps(box, steps, warmup):
    import torch
    box.step(warmup)
    box.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    box.step(steps)
    box.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def copy_ceiling_gbs():
    """device-to-device copy of 1 GiB, timed with torch events (read + write bytes)."""
    import torch
    n = 1 << 28  # floats: 1 GiB per buffer
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def err_dt(n):
    """fixed step of the error runs: about the CFL 0.5 step of the workload (256^3: 1.8e-3)"""
    return 0.002 if n <= 256 else 0.002 * 256 / n * 0.85


def run_fixed(n, prec, steps, state_precision=0, seed=0):
    dt = err_dt(n)
    b = make_box(n, prec, adaptive=False, dt=dt, state_precision=state_precision)
    b.init_random(seed, 0, 4.0, 0.5)
    b.rhs_int_fact(dt)
    b.step(steps)
    out = b.diagnostics()
    b.close()
    return out


def errors(ref, got):
    import numpy as np
    (d64, s64), (d32, s32) = ref, got
    if not (np.all(np.isfinite(s64)) and np.all(np.isfinite(s32))):
        return {"spectrum_max_shell": None, "spectrum_l2": None, "energy": None, "note": "non-finite run"}
    nz = s64 > 0
    return {"spectrum_max_shell": float(np.max(np.abs(s32[nz] - s64[nz]) / s64[nz])),
            "spectrum_l2": float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64)),
            "energy": float(abs(d32[0] - d64[0]) / d64[0])}


def large_form(n, steps, err_steps):
    """The precision form at a larger size (nu = 0.002): steps/s of the FP32 form with the FP64
    state and of FP64 (adaptive dt, CFL 0.5), and the error of the former against the latter after
    err_steps fixed-dt steps from the same field."""
    global NU
    saved, NU = NU, 0.002
    try:
        rates = {}
        for key, prec, sp, k in (("fp32_state64", 32, 64, steps), ("fp64", 64, 0, max(1, steps // 2))):
            b = make_box(n, prec, state_precision=sp)
            b.init_random(0, 0, 4.0, 0.5)
            rates[key] = k / timed_steps(b, k, 10)
            b.close()
        out = {"n": n, "nu": NU, "steps_per_s": rates["fp32_state64"], "fp64_steps_per_s": rates["fp64"],
               "speedup_vs_fp64": rates["fp32_state64"] / rates["fp64"]}
        if err_steps > 0:
            ref = run_fixed(n, 64, err_steps)
            out["err_vs_fp64"] = errors(ref, run_fixed(n, 32, err_steps, state_precision=64))
            out["err_steps"], out["err_dt"] = err_steps, err_dt(n)
        return out
    finally:
        NU = saved


def kernel_table(box, reps):
    """HIP-event time and algorithmic GB/s of each step pass of `box`."""
    from dnsbox_amd import _lib
    ms = box.time_passes(reps)
    byts = box.pass_bytes()
    out = {}
    for name, per_step in zip(_lib.PASS_NAMES, _lib.PASS_PER_STEP):
        out[name] = {"ms": ms[name], "alg_bytes": byts[name],
                     "gbs": byts[name] / (ms[name] * 1e-3) / 1e9, "ms_per_step": ms[name] * per_step}
    return out


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` for `workload` from the rocprofv3 counter summary on
    file (lines `workload kernel fetch_bytes write_bytes`), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[0] == workload and p[1] == kernel and re.fullmatch(r"[0-9.eE+]+", p[2]):
                return float(p[2]) + float(p[3])
    return None


def adjoint_rate(reps=3, descent_iters=50):
    """64^3 FP64 residual gradients per second (T = 200, dt = 0.005, checkpoint every 20)
    and a central-difference check of one gradient along a seeded solenoidal direction."""
    import torch
    from dnsbox_amd import Box
    from dnsbox_amd.adjoint import adjoint_descent, residual
    n, T, dt, k = 64, 200, 0.005, 20
    with Box(n=n, precision=64, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.init_random(0, 0, 4.0, 0.5)
        u0 = b.state_tensor()
        b.init_random(1, 0, 4.0, 0.5)
        v = b.state_tensor()

        def grad():
            x = u0.clone().requires_grad_(True)
            J = residual(b, x, T, dt, k)
            (g,) = torch.autograd.grad(J, x)
            return J, g

        for _ in range(2):   # the first gradient also creates the twin box
            J, g = grad()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            J, g = grad()
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / reps
        eps = 1e-6
        with torch.no_grad():
            jp = float(residual(b, u0 + eps * v, T, dt, k))
            jm = float(residual(b, u0 - eps * v, T, dt, k))
        fd = (jp - jm) / (2 * eps)
        an = float((g.conj() * v).real.sum())
        # the configured descent: 50 iterations with the FP64 backtracking line search
        hist, t_desc = [float(J.detach())], 0.0
        if descent_iters > 0:
            t0 = time.perf_counter()
            _, hist = adjoint_descent(b, u0, T, dt, k=k, iters=descent_iters)
            torch.cuda.synchronize()
            t_desc = time.perf_counter() - t0
        from dnsbox_amd.adjoint import _mixed_fits, _segment_length
        seg = _segment_length(T, k, u0.numel() * u0.element_size(), u0.device, "auto")
        mixed = _mixed_fits(b, T - 1, u0.device)
    gib = 2 * (T - 1) * u0.numel() * u0.element_size() / 2**30
    backward = (f"trajectory recorded by the forward ({gib:.2f} GiB of states and predicted states"
                + (f", {3 * gib:.2f} GiB of x-pass inputs read by the step adjoints)" if mixed else ")")
                if seg == T else f"checkpoint every {k} steps, segments recomputed on a second stream")
    return {"grads_per_s": 1.0 / secs, "ms_per_grad": secs * 1e3, "n": n, "T": T, "dt": dt,
            "checkpoint_every": k, "backward": backward, "dtype": "f64", "J": float(J.detach()),
            "fd_rel_err": abs(fd - an) / abs(fd),
            "descent": {"iterations": len(hist) - 1, "J0": hist[0], "J_final": hist[-1],
                        "monotone": all(b_ <= a_ for a_, b_ in zip(hist, hist[1:])), "seconds": t_desc},
            "workload": "64^3 FP64, nu=0.02, f=(sin 4y,0,0), J=1/2<|Phi_T(u)-u|^2>"}


def batched_rates(steps=500, T=100, k=20, reps=3, warm=2):
    """The 48-box group: forward box-steps/s (hipGraph-replayed step of the boxes' own states)
    and forward + adjoint box-steps/s (checkpointed gradient of the T-step map)."""
    import torch
    from dnsbox_amd import BoxGroup
    from dnsbox_amd.adjoint import group_residual_grads
    from dnsbox_amd.workloads import batched48, init_batched48
    dt = 0.005
    params, k0s = batched48(32, dt=dt)
    sizes = [p["n"] for p in params]
    with BoxGroup(params) as g:
        init_batched48(g.boxes, k0s)
        us = []
        for b in g.boxes:
            b.rhs_int_fact(dt)
            us.append(b.state_tensor())
        g.step(10)
        g.sync()
        t0 = time.perf_counter()
        g.step(steps)
        g.sync()
        fwd = len(sizes) * steps / (time.perf_counter() - t0)
        # algorithmic HBM bytes of one group step: every box's middle passes twice and its two
        # fused z passes (the pipeline-start z-inverse aside), as each box's next step runs them
        step_bytes = 0.0
        for b in g.boxes:
            pb = b.pass_bytes()
            step_bytes += 2 * (pb["yinv6"] + pb["xcross"] + pb["yfwd3"]) + pb["zpred_zinv"] + pb["zcorr_zinv"]
        dts = [dt] * len(sizes)
        for _ in range(warm):   # the first gradients create the twin group and the descriptor sets
            group_residual_grads(g, us, T, dts, k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            group_residual_grads(g, us, T, dts, k)
        torch.cuda.synchronize()
        secs = (time.perf_counter() - t0) / reps
    group_steps_per_s = fwd / len(sizes)
    achieved = step_bytes * group_steps_per_s / 1e9
    return {"boxes": len(sizes), "sizes": "16x32^3 + 16x48^3 + 16x64^3", "dtype": "f32",
            "forward_box_steps_per_s": fwd,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_group_step": step_bytes,
                         "note": "whole group step (all passes of the 48 boxes), algorithmic bytes / step time"},
            "fwd_adj_box_steps_per_s": len(sizes) * T / secs,
            "ms_per_batched_gradient": secs * 1e3, "T": T, "checkpoint_every": k, "dt": dt}


def grid_rates(steps=200, reps=10):
    """Steps/s of non-cubic grids (CFL 0.5, nu = 0.005, f = sin 4y): 256 x 256 x 128 in the FP32
    form with the FP64 state and in FP64, and the 2 pi x 2 pi x 4 pi box at 64 x 64 x 128
    (alpha_z = 1/2) in FP64, with the per-pass times and algorithmic rates of each step."""
    from dnsbox_amd import Box
    out = {}
    for key, (nx, ny, nz), az, prec, sp, k in (
            ("256x256x128_fp32_state64", (256, 256, 128), 1.0, 32, 64, steps),
            ("256x256x128_fp64", (256, 256, 128), 1.0, 64, 0, max(1, steps // 2)),
            ("64x64x128_fp64_lz_4pi", (64, 64, 128), 0.5, 64, 0, steps)):
        b = Box(n=nx, nx=nx, ny=ny, nz=nz, alpha_z=az, precision=prec, state_precision=sp, nu=NU,
                kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=1, cfl_min=0.5, cfl_max=0.5, dt_max=0.01,
                dt=0.002)
        b.init_random(0, 0, 4.0, 0.5)
        rate = k / timed_steps(b, k, 10)
        kt = kernel_table(b, reps)
        b.close()
        out[key] = {"steps_per_s": rate, "point_steps_per_s": rate * nx * ny * nz,
                    "kernels": {q: {"ms": round(v["ms"], 5), "gbs": round(v["gbs"], 1)} for q, v in kt.items()}}
    return out


def cpu_baseline(n, threads, nsteps=2):
    """`nsteps` IF-Heun steps of the FP64 CPU oracle (reference-structure restatement: 15 3-D
    transforms per RHS) at n^3 on `threads` host threads; about 10 s at 256^3 on 16 threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(n, nu=NU, kolmo_amp=1.0, kolmo_k=4, nthreads=threads)
    f = o.init_solenoidal(0, 0, 4.0, 0.5)
    t0 = time.perf_counter()
    o.heun(f, nsteps, 0.002)
    dt = time.perf_counter() - t0
    return nsteps / dt, dt, nsteps


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the step of one box, or a group of "
                          "boxes, per GPU; --gpus N > 1 is not supported (sharding is design-only)"}))
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    global NU
    NU = args.nu if args.nu is not None else (0.005 if n <= 256 else 0.002)
    log(f"n={n} steps={args.steps} warmup={args.warmup}")

    # ---- headline: FP32 transforms + FP64 state, timed ----
    bh = make_box(n, 32, state_precision=64)
    bh.init_random(0, 0, 4.0, 0.5)
    secs = timed_steps(bh, args.steps, args.warmup)
    value = args.steps / secs
    t_end, dt_end = bh.time()
    log(f"fp32 (fp64 state) {value:.1f} steps/s (t={t_end:.4f}, dt={dt_end:.3e})")
    pruned = bh.pruning()
    kernels = kernel_table(bh, args.reps)
    bh.close()

    # ---- the same form with the transforms unpruned (same results to rounding) ----
    unpruned = None
    if args.unpruned_steps > 0:
        bu = make_box(n, 32, state_precision=64)
        bu.init_random(0, 0, 4.0, 0.5)
        bu.set_pruning(False)
        unpruned = args.unpruned_steps / timed_steps(bu, args.unpruned_steps, args.warmup)
        bu.close()
        log(f"fp32 (fp64 state), unpruned transforms: {unpruned:.1f} steps/s")
    dom = max(kernels, key=lambda q: kernels[q]["ms_per_step"])
    copy_gbs = copy_ceiling_gbs()
    log(f"passes {json.dumps({q: round(v['gbs']) for q, v in kernels.items()})} copy={copy_gbs:.0f} GB/s")

    # ---- all-FP32 form ----
    pure = None
    if args.pure_steps > 0:
        bp = make_box(n, 32)
        bp.init_random(0, 0, 4.0, 0.5)
        pure_rate = args.pure_steps / timed_steps(bp, args.pure_steps, args.warmup)
        pure_kernels = kernel_table(bp, args.reps)
        bp.close()
        pure = {"steps_per_s": pure_rate,
                "kernels": {q: {"ms": round(v["ms"], 5), "gbs": round(v["gbs"], 1)} for q, v in pure_kernels.items()}}
        log(f"fp32 (fp32 state) {pure_rate:.1f} steps/s")

    # ---- FP64 ----
    b64 = make_box(n, 64)
    b64.init_random(0, 0, 4.0, 0.5)
    fp64_rate = args.fp64_steps / timed_steps(b64, args.fp64_steps, args.warmup)
    b64.close()
    log(f"fp64 {fp64_rate:.1f} steps/s")
    # the FP64 step with whole-field transforms, as the reference transforms (p3dfft)
    fp64_full = None
    if args.unpruned_steps > 0:
        b64 = make_box(n, 64)
        b64.init_random(0, 0, 4.0, 0.5)
        b64.set_pruning(False)
        fp64_full = min(args.unpruned_steps, args.fp64_steps) / timed_steps(
            b64, min(args.unpruned_steps, args.fp64_steps), args.warmup)
        b64.close()
        log(f"fp64, unpruned transforms {fp64_full:.1f} steps/s")

    # ---- errors vs FP64 (same field, same fixed dt) ----
    err = err_pure = None
    if args.err_steps > 0:
        ref = run_fixed(n, 64, args.err_steps)
        err = errors(ref, run_fixed(n, 32, args.err_steps, state_precision=64))
        if pure is not None:
            err_pure = errors(ref, run_fixed(n, 32, args.err_steps))
            pure["err_vs_fp64"] = err_pure
        log(f"vs fp64 after {args.err_steps} steps: fp64-state {err}; all-fp32 {err_pure}")

    large = None
    if args.large_n > 0 and args.large_n != n:
        large = large_form(args.large_n, args.large_steps, args.large_err_steps)
        log(f"{args.large_n}^3 (nu 0.002): fp32 (fp64 state) {large['steps_per_s']:.1f} steps/s, fp64 "
            f"{large['fp64_steps_per_s']:.1f}; vs fp64 {large.get('err_vs_fp64')}")

    adj = adjoint_rate(descent_iters=args.descent_iters) if args.adjoint else None
    if adj:
        log(f"adjoint {adj['grads_per_s']:.2f} grads/s, fd rel err {adj['fd_rel_err']:.2e}")
    bat = batched_rates() if args.batched else None
    grids = grid_rates() if args.grids else None
    if grids:
        log("grids " + ", ".join(f"{k} {v['steps_per_s']:.1f} steps/s" for k, v in grids.items()))
    if bat:
        log(f"batched {bat['forward_box_steps_per_s']:.0f} box-steps/s forward, "
            f"{bat['fwd_adj_box_steps_per_s']:.0f} forward+adjoint")

    # ---- CPU baseline (oracle, restatement of the reference structure) ----
    cpu = None
    if args.cpu_baseline:
        host_cpus = os.cpu_count() or 1
        usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else host_cpus
        threads = min(args.cpu_threads, usable)
        rate, secs_cpu, cpu_steps = cpu_baseline(n, threads)
        cpu = {"value": rate, "unit": "time-steps/s", "cores": threads, "kind": "port",
               "sample": f"{cpu_steps} IF-Heun steps of the FP64 CPU oracle at {n}^3 (15 transforms/RHS, "
                         f"OpenMP), {secs_cpu:.1f} s, on {threads} threads of a host with {host_cpus} "
                         f"logical CPUs ({usable} usable by this process)"}
        log(f"cpu {rate:.4f} steps/s on {threads} threads")

    kd = kernels[dom]
    workload = f"{n}cubed_fp32_state64"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded solenoidal field, E(k)~k^4 exp(-2(k/4)^2), E=0.5)",
        "config": {"workload": f"{n}^3 forced box, IF-Heun, CFL 0.5, nu={NU}, f=(sin 4y,0,0), "
                               "FP32 FFT/nonlinear, FP64 state/IF/CFL/diagnostics",
                   "n": n, "nu": NU, "cfl": 0.5, "stepper": "IF-Heun",
                   "transforms": "pruned to the 2/3-rule cube" if pruned else "full"},
        "unpruned_steps_per_s": unpruned,
        "err_vs_fp64": err,
        "err_steps": args.err_steps,
        "err_dt": err_dt(n),
        "fp64_steps_per_s": fp64_rate,
        "speedup_vs_fp64": (value / fp64_rate) if fp64_rate else None,
        "fp64_unpruned_steps_per_s": fp64_full,
        "speedup_vs_fp64_unpruned": (value / fp64_full) if fp64_full else None,
        "fp32_pure": pure,
        "speedup_fp32_pure_vs_fp64": (pure["steps_per_s"] / fp64_rate) if pure and fp64_rate else None,
        "roofline": {"bound": "hbm", "achieved": kd["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kd["gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(dom, workload),
                     "kernel": dom, "alg_bytes_per_launch": kd["alg_bytes"], "ms_per_launch": kd["ms"],
                     "copy_ceiling_gbs": copy_gbs,
                     "note": ("the x pass is bound by VALU issue and latency at 3 waves per SIMD (about 40 M "
                              "wave64 VALU, i.e. 66 us of issue, and 4.3 M LDS instructions per launch at "
                              "256^3 FP32), not by HBM"
                              if dom == "xcross" else None)},
        "kernels": {q: {"ms": round(v["ms"], 5), "gbs": round(v["gbs"], 1)} for q, v in kernels.items()},
        "large_box": large,
        "adjoint": adj,
        "batched": bat,
        "grids": grids,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
