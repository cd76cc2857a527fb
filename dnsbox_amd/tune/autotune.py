"""Autotuner of the FFT radices and tiles of the step passes (dnsbox_amd/csrc/tuned_cfg.hpp).

For one box size and precision it varies one axis of the configuration at a time around the
current table entry:
  z  : tile width and symmetric radix plan of the z passes   (timed: zpred_zinv + zcorr_zinv)
  y  : tile width and radix plan of the y passes             (timed: 2 yinv6 + 2 yfwd3)
  x  : lines per block and row radix plan of the x pass, with the LDS row layout chosen by
       the bank model (ldsmodel.py)                           (timed: 2 xcross)
Every candidate is compiled into the timing harness (csrc/tune_main.hip, one executable per
candidate, compiled in parallel on the host) and timed on the GPU inside real IF-Heun steps;
candidates whose energy after five fixed-dt steps differs from the table entry's beyond
rounding are rejected.  The best of each axis are combined and written into the table.

    python -m dnsbox_amd.tune.autotune build --n 256 --prec 32 --state 64   # host: compile
    python -m dnsbox_amd.tune.autotune run                                   # GPU: time
    python -m dnsbox_amd.tune.autotune emit                                  # host: update table
    python -m dnsbox_amd.tune.autotune all --n 256 --prec 32 --state 64     # all three, GPU host

Sizes that run in box groups (N <= 128) keep 256-thread blocks in every pass.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import itertools
import json
import os
import subprocess
import sys
import time

from .cfgtable import CSRC, GROUP_SIZES, Cfg, emit, parse
from .ldsmodel import best_layouts

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WORK = os.path.join(PKG, "_tune")
HIPCC = os.environ.get("HIPCC", "hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
LDS_MAX = 160 * 1024 - 4096          # bytes of LDS per block, less the twiddles and scratch words
THREADS_BIG = (256, 512, 1024)


def _plans(length, radices, max_stages=4):
    out = set()
    for s in range(1, max_stages + 1):
        for p in itertools.product(radices, repeat=s):
            prod = 1
            for r in p:
                prod *= r
            if prod != length:
                continue
            rm = max(p)
            if all(rm % r == 0 for r in p):
                out.add(p)
    return sorted(out, key=lambda p: (len(p), -max(p), p))


def _radices(n):
    return (2, 3, 4, 8, 12, 16) if n % 3 == 0 else (2, 4, 8, 16)


def _threads_ok(n, threads):
    return threads == 256 if n in GROUP_SIZES else threads in THREADS_BIG


def z_candidates(n, esz, base, limit=8):
    out = []
    for zp in _plans(n, _radices(n), 3):
        if zp[0] != zp[-1]:
            continue
        rm = max(zp)
        for zw in (4, 8, 16, 32, 64):
            if not _threads_ok(n, zw * (n // rm)) or n * zw * esz > LDS_MAX:
                continue
            if rm * esz * 9 > 1024 * 2:     # registers: 9 fields of rm elements live in the fused pass
                continue
            out.append(base.with_(zw=zw, zp=zp))
    out.sort(key=lambda c: (c != base, len(c.zp), abs(max(c.zp) - 8), -c.zw))
    return out[:limit]


def y_candidates(n, esz, base, limit=8):
    out = []
    for yp in _plans(n, _radices(n), 3):
        rm = max(yp)
        for yw in (16, 32, 64):
            if not _threads_ok(n, yw * (n // rm)) or n * yw * esz > LDS_MAX:
                continue
            if rm * esz > 1024:
                continue
            out.append(base.with_(yw=yw, yp=yp))
    out.sort(key=lambda c: (c != base, len(c.yp), abs(max(c.yp) - 16), -c.yw))
    return out[:limit]


def x_candidates(n, prec, base, limit=8):
    esz = 8 if prec == 32 else 16
    h = n // 2 + 1
    rows = n * n
    out = [base]
    plans = _plans(n // 2, _radices(n // 2), 3)
    plans = [p for p in plans if len(p) == min(len(q) for q in plans)][:3]
    for rp in plans:
        for ntx in ((256,) if n in GROUP_SIZES else (128, 256, 512)):
            for rb in (1, 2, 4, 8, 16, 32, 64):
                # 2 to 16 input loads per thread
                if rows % rb or rb * h > ntx * 16 or (rb * h < 2 * ntx and rb * 2 <= rows // 256):
                    continue
                lay = best_layouts(n, prec, rb, ntx, rp, top=1, periods=(0, 16, 32), extra=8)
                xpd, xpad, hp, _ = lay[0]
                if 12 * rb * hp * esz > LDS_MAX or 6 * rb * hp * esz > 64 * 1024:
                    continue
                # the bank model's best layout, and one pad every 64 elements (measured faster
                # than the model's choice for 256^3 FP32, slower elsewhere)
                for xd, xa in ((xpd, xpad), (64, 1)):
                    c = base.with_(rb=rb, ntx=ntx, rp=rp, xpd=xd, xpad=xa)
                    if c not in out:
                        out.append(c)
    out.sort(key=lambda c: (c != base, abs(c.rb * h * 256 // c.ntx - 4 * 129), len(c.rp)))
    return out[:limit]


def _tag(n, prec, state):
    return f"n{n}_f{prec}_s{state}"


def build(n, prec, state, jobs=8, limit=8):
    """Compile the harness for every candidate; returns the manifest path."""
    table = parse()
    tname = "float" if prec == 32 else "double"
    base = table[(n, tname)]
    esz = 8 if prec == 32 else 16
    cands = [("base", base)]
    for axis, lst in (("z", z_candidates(n, esz, base, limit)), ("y", y_candidates(n, esz, base, limit)),
                      ("x", x_candidates(n, prec, base, limit))):
        cands += [(axis, c) for c in lst if c != base]
    out_dir = os.path.join(WORK, _tag(n, prec, state))
    os.makedirs(out_dir, exist_ok=True)
    sname = "double" if state == 64 else tname

    def one(i):
        axis, c = cands[i]
        exe = os.path.join(out_dir, f"cand{i:03d}")
        cmd = [HIPCC] + FLAGS + [f"-DDNSBOX_TUNE_N={n}", f"-DDNSBOX_TUNE_T={tname}", f"-DDNSBOX_TUNE_S={sname}",
                                 f"-DDNSBOX_TUNE_CFG={c.cxx()}", os.path.join(CSRC, "tune_main.hip"), "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        ok = r.returncode == 0
        if not ok:
            sys.stderr.write(f"[autotune] candidate {i} ({axis} {c.cxx()}) does not compile:\n"
                             f"{r.stderr[-2000:]}\n")
        return {"id": i, "axis": axis, "cfg": c.__dict__, "exe": os.path.relpath(exe, PKG) if ok else None}

    t0 = time.time()
    with cf.ThreadPoolExecutor(jobs) as ex:
        entries = list(ex.map(one, range(len(cands))))
    manifest = {"n": n, "prec": prec, "state": state, "built": time.strftime("%Y-%m-%d %H:%M:%S"),
                "candidates": entries}
    path = os.path.join(out_dir, "manifest.json")
    with open(path, "w") as f:
        json.dump(manifest, f, indent=1)
    print(f"[autotune] {sum(e['exe'] is not None for e in entries)}/{len(entries)} candidates built "
          f"for {_tag(n, prec, state)} in {time.time() - t0:.0f} s -> {path}", flush=True)
    return path


def run(manifests, reps=40, timeout=120, out=None, rounds=2):
    """Time every built candidate on the GPU (one process each), `rounds` times in turn;
    each pass keeps its fastest time over the rounds."""
    best = {}
    for rnd in range(rounds):
        for mpath in manifests:
            with open(mpath) as f:
                man = json.load(f)
            for e in man["candidates"]:
                if not e["exe"]:
                    continue
                key = (man["n"], man["prec"], man["state"], e["id"])
                if rnd > 0 and key not in best:
                    continue   # failed in the first round
                exe = os.path.join(PKG, e["exe"])
                print(f"[autotune] round {rnd} {_tag(man['n'], man['prec'], man['state'])} candidate {e['id']} "
                      f"({e['axis']})", flush=True)
                try:
                    r = subprocess.run([exe, str(reps)], capture_output=True, text=True, timeout=timeout)
                except subprocess.TimeoutExpired:
                    print("[autotune]   timed out", flush=True)
                    best.pop(key, None)
                    continue
                if r.returncode != 0:
                    print(f"[autotune]   failed: {r.stderr.strip()[-300:]}", flush=True)
                    best.pop(key, None)
                    continue
                t = json.loads(r.stdout.strip().splitlines()[-1])
                print(f"[autotune]   {json.dumps(t)}", flush=True)
                if key in best:
                    old = best[key]["t"]
                    best[key]["t"] = {k: (v if k == "ekin" else min(v, old[k])) for k, v in t.items()}
                else:
                    best[key] = {"n": man["n"], "prec": man["prec"], "state": man["state"], "id": e["id"],
                                 "axis": e["axis"], "cfg": e["cfg"], "t": t}
    results = list(best.values())
    out = out or os.path.join(WORK, "results.json")
    with open(out, "w") as f:
        json.dump(results, f, indent=1)
    print(f"[autotune] {len(results)} timings -> {out}", flush=True)
    return out


AXIS_COST = {"z": lambda t: t["zpred_zinv"] + t["zcorr_zinv"],
             "y": lambda t: 2 * (t["yinv6"] + t["yfwd3"]),
             "x": lambda t: 2 * t["xcross"]}
AXIS_KEYS = {"z": ("zw", "zp"), "y": ("yw", "yp"), "x": ("rb", "ntx", "rp", "xpd", "xpad")}


def choose(results):
    """Per (n, prec): the best candidate of each axis (energy within rounding of the table's)."""
    picks = {}
    for key in sorted({(r["n"], r["prec"], r["state"]) for r in results}):
        rs = [r for r in results if (r["n"], r["prec"], r["state"]) == key]
        base = [r for r in rs if r["axis"] == "base"]
        if not base:
            continue
        b = base[0]
        tol = 1e-5 if key[1] == 32 else 1e-11
        ek = b["t"]["ekin"]
        good = [r for r in rs if abs(r["t"]["ekin"] - ek) <= tol * abs(ek)]
        cfg = dict(b["cfg"])
        notes = [f"N={key[0]} FP{key[1]} (state FP{key[2]}): table entry step {b['t']['step']:.4f} ms"]
        for axis in ("z", "y", "x"):
            pool = [r for r in good if r["axis"] in (axis, "base")]
            best = min(pool, key=lambda r: AXIS_COST[axis](r["t"]))
            for k in AXIS_KEYS[axis]:
                cfg[k] = best["cfg"][k]
            notes.append(f"  {axis}: {AXIS_COST[axis](best['t']):.4f} ms/step "
                         f"(table {AXIS_COST[axis](b['t']):.4f}, {len(pool)} candidates)")
        rejected = len(rs) - len(good)
        if rejected:
            notes.append(f"  {rejected} candidate(s) rejected: energy differs from the table entry's")
        picks[key] = (cfg, notes)
    return picks


def emit_results(results_path):
    with open(results_path) as f:
        results = json.load(f)
    table = parse()
    notes = []
    for (n, prec, state), (cfg, nn) in choose(results).items():
        c = Cfg(**{k: tuple(v) if isinstance(v, list) else v for k, v in cfg.items()})
        table[(n, "float" if prec == 32 else "double")] = c
        notes += nn + [f"  -> {c.cxx()}"]
    notes.append(time.strftime("tuned %Y-%m-%d"))
    emit(table, notes)
    print("\n".join(notes))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("cmd", choices=("candidates", "build", "run", "emit", "all"))
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--prec", type=int, default=32)
    ap.add_argument("--state", type=int, default=0, help="state precision (0 = prec)")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--limit", type=int, default=8, help="candidates per axis")
    ap.add_argument("--reps", type=int, default=40)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--results", default=os.path.join(WORK, "results.json"))
    ap.add_argument("--manifest", action="append", default=None)
    a = ap.parse_args(argv)
    state = a.state or a.prec
    if a.cmd == "candidates":
        table = parse()
        base = table[(a.n, "float" if a.prec == 32 else "double")]
        esz = 8 if a.prec == 32 else 16
        for axis, lst in (("z", z_candidates(a.n, esz, base, a.limit)), ("y", y_candidates(a.n, esz, base, a.limit)),
                          ("x", x_candidates(a.n, a.prec, base, a.limit))):
            for c in lst:
                print(axis, c.cxx())
        return
    if a.cmd in ("build", "all"):
        m = build(a.n, a.prec, state, a.jobs, a.limit)
        a.manifest = (a.manifest or []) + [m]
    if a.cmd in ("run", "all"):
        mans = a.manifest or [os.path.join(WORK, d, "manifest.json") for d in sorted(os.listdir(WORK))
                              if os.path.exists(os.path.join(WORK, d, "manifest.json"))]
        run(mans, a.reps, out=a.results, rounds=a.rounds)
    if a.cmd in ("emit", "all"):
        emit_results(a.results)


if __name__ == "__main__":
    main()
