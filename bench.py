#!/usr/bin/env python3
"""Benchmark of the MI355X pseudospectral Navier-Stokes step (one GPU).

Workload (the headline metric's configuration): 256^3 box, FP32 FFTs and nonlinear
term with FP64 integrating factor / CFL control / diagnostics, nu = 0.005,
Kolmogorov forcing f = (sin 4y, 0, 0), IF-Heun predictor-corrector with the CFL
controller at Courant 0.5, seeded solenoidal initial field E(k) ~ k^4 exp(-2 (k/4)^2)
normalised to E = 0.5.  A step is one IF-Heun step (two RHS evaluations) of the
whole box, with the state resident in HBM.

Printed as one JSON line: steps/s (FP32), the FP64 step rate beside it, the FP32
spectrum/energy error against the FP64 run after the same number of fixed-dt steps,
the HBM roofline fraction of the dominant kernel (HIP-event timed), and the CPU
oracle timed on the host cores (a reported baseline, not the target).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-steps/s at 256³ FP32 vs FP64 (err stated), % HBM roofline, "
          "adjoint grads/s")
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--err-steps", type=int, default=1000,
                    help="fixed-dt steps of the FP32-vs-FP64 error run (0 = skip)")
    ap.add_argument("--fp64-steps", type=int, default=200, help="timed FP64 steps")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle (1/0)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--reps", type=int, default=20, help="launches per pass for kernel timing")
    return ap.parse_args()


def make_box(n, precision, adaptive=True, dt=0.002, state_precision=0):
    from dnsbox_amd import Box
    return Box(n=n, precision=precision, nu=0.005, kolmo_amp=1.0, kolmo_k=4, bandlim=0,
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
    """device-to-device copy of 2 GiB, timed with torch events (read + write bytes)."""
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


def run_fixed(n, prec, steps, state_precision=0, seed=0):
    b = make_box(n, prec, adaptive=False, dt=0.002, state_precision=state_precision)
    b.init_random(seed, 0, 4.0, 0.5)
    b.rhs_int_fact(0.002)
    b.step(steps)
    out = b.diagnostics()
    b.close()
    return out


def errors(ref, got):
    import numpy as np
    (d64, s64), (d32, s32) = ref, got
    nz = s64 > 0
    return {"spectrum_max_shell": float(np.max(np.abs(s32[nz] - s64[nz]) / s64[nz])),
            "spectrum_l2": float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64)),
            "energy": float(abs(d32[0] - d64[0]) / d64[0])}


def cpu_baseline(n, threads):
    """One IF-Heun step of the FP64 CPU oracle (reference-structure restatement: 15 3-D
    transforms per RHS) at n^3 on `threads` host threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    o = Oracle(n, nu=0.005, kolmo_amp=1.0, kolmo_k=4, nthreads=threads)
    f = o.init_solenoidal(0, 0, 4.0, 0.5)
    t0 = time.perf_counter()
    o.heun(f, 1, 0.002)
    dt = time.perf_counter() - t0
    return 1.0 / dt, dt


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one box per GPU on a single GPU; "
                          "--gpus N > 1 is not supported (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    import numpy as np
    import torch
    from dnsbox_amd import _lib

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.n
    print(f"[bench] n={n} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)

    # ---- FP32 timed run (the metric) ----
    b32 = make_box(n, 32)
    b32.init_random(0, 0, 4.0, 0.5)
    secs = timed_steps(b32, args.steps, args.warmup)
    value = args.steps / secs
    t32, dt32 = b32.time()
    print(f"[bench] fp32 {value:.1f} steps/s (t={t32:.4f}, dt={dt32:.3e})", file=sys.stderr, flush=True)

    # ---- kernel passes, HIP events on the box stream ----
    ms = b32.time_passes(args.reps)
    fs_bytes = n * n * (n // 2 + 1) * 8  # one FP32 complex spectral field
    per_step = {"zinv_curl": 0, "yinv6": 2, "xcross": 2, "yfwd3": 2, "zpred_zinv": 1, "zcorr_zinv": 1}
    kernels = {}
    for name, (r, w) in zip(_lib.PASS_NAMES, _lib.PASS_FIELDS):
        byts = (r + w) * fs_bytes
        kernels[name] = {"ms": ms[name], "alg_bytes": byts,
                         "gbs": byts / (ms[name] * 1e-3) / 1e9,
                         "ms_per_step": ms[name] * per_step[name]}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    copy_gbs = copy_ceiling_gbs()
    b32.close()
    print(f"[bench] passes {json.dumps({k: round(v['gbs']) for k, v in kernels.items()})} "
          f"copy={copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)

    # ---- FP64 timed run (same workload in FP64) ----
    b64 = make_box(n, 64)
    b64.init_random(0, 0, 4.0, 0.5)
    secs64 = timed_steps(b64, args.fp64_steps, args.warmup)
    fp64_rate = args.fp64_steps / secs64
    b64.close()
    print(f"[bench] fp64 {fp64_rate:.1f} steps/s", file=sys.stderr, flush=True)

    # ---- FP32 state-in-FP64 variant: rate ----
    b3 = make_box(n, 32, state_precision=64)
    b3.init_random(0, 0, 4.0, 0.5)
    secs3 = timed_steps(b3, args.steps, args.warmup)
    rate_s64 = args.steps / secs3
    b3.close()
    print(f"[bench] fp32 + fp64 state {rate_s64:.1f} steps/s", file=sys.stderr, flush=True)

    # ---- FP32 error vs FP64 (same field, same fixed dt, err_steps steps) ----
    err32 = err_s64 = None
    if args.err_steps > 0:
        ref = run_fixed(n, 64, args.err_steps)
        err32 = errors(ref, run_fixed(n, 32, args.err_steps))
        err_s64 = errors(ref, run_fixed(n, 32, args.err_steps, state_precision=64))
        print(f"[bench] vs fp64 after {args.err_steps} steps: fp32 {err32}; fp32+fp64 state {err_s64}",
              file=sys.stderr, flush=True)

    # ---- CPU baseline (oracle, port of the reference structure) ----
    cpu = None
    if args.cpu_baseline:
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        rate, secs_cpu = cpu_baseline(n, threads)
        cpu = {"value": rate, "unit": "time-steps/s", "cores": threads, "kind": "port",
               "sample": f"1 IF-Heun step of the FP64 CPU oracle at {n}^3 (15 transforms/RHS, "
                         f"OpenMP), {secs_cpu:.1f} s"}
        print(f"[bench] cpu {rate:.4f} steps/s on {threads} threads", file=sys.stderr, flush=True)

    k = kernels[dom]
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
        "config": {"workload": f"{n}^3 forced box, IF-Heun, CFL 0.5, nu=0.005, f=(sin 4y,0,0), "
                               "FP32 FFT/nonlinear + FP64 IF/CFL/diagnostics",
                   "n": n, "nu": 0.005, "cfl": 0.5, "stepper": "IF-Heun"},
        "fp64_steps_per_s": fp64_rate,
        "speedup_fp32_vs_fp64": value / fp64_rate,
        "fp32_err_vs_fp64": err32,
        "fp32_state64_steps_per_s": rate_s64,
        "fp32_state64_err_vs_fp64": err_s64,
        "speedup_fp32_state64_vs_fp64": rate_s64 / fp64_rate,
        "err_steps": args.err_steps,
        "roofline": {"bound": "hbm", "achieved": k["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": k["gbs"] / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "copy_ceiling_gbs": copy_gbs},
        "kernels": {kk: {"ms": round(v["ms"], 5), "gbs": round(v["gbs"], 1)} for kk, v in kernels.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
