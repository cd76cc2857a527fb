"""Writes tests/golden/forced128_fp64_adaptive.json: the oracle's trajectory of the 128^3 FP64
forced box over 1000 IF-Heun steps with the CFL 0.5 controller, diagnostics every 100 steps.

Workload: seeded solenoidal field (seed 0, box 0, E(k) ~ k^4 exp(-2 (k/4)^2), E = 0.5),
nu = 0.01, f = (sin 4y, 0, 0), dt_0 = 0.002, Courant band [0.5, 0.5], dt_max = 0.01; the step
size is set from the current field before every step (dnsboxheun.f90:153-157,227-246 with the
Courant number io.f90:664-676 of the state being stepped).

The oracle (oracle/dnsbox_oracle.c) is run here, on the CPU (about 1 s per step at 128^3 on 8
threads); the GPU test compares against the committed file and runs no oracle.

    python tests/golden/make_forced128_fixture.py [--threads 8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import Oracle  # noqa: E402

OUT = os.path.join(HERE, "forced128_fp64_adaptive.json")
N, NU, DT0, CFL, DTMAX, STEPS, EVERY = 128, 0.01, 0.002, 0.5, 0.01, 1000, 100


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    o = Oracle(N, nu=NU, kolmo_amp=1.0, kolmo_k=4, nthreads=args.threads)
    u = o.init_solenoidal(0, 0, 4.0, 0.5)
    dt, t = DT0, 0.0
    rows = []
    t0 = time.time()
    for chk in range(EVERY, STEPS + 1, EVERY):
        u, dt, t = o.heun_adaptive(u, EVERY, dt, CFL, CFL, DTMAX, t)
        diag, spec = o.diagnostics(u, dt)
        rows.append({"step": chk, "t": t, "dt": dt, "Ekin": diag[0], "Disp": diag[1], "Prod": diag[2],
                     "Courant": diag[3], "divMax": diag[4],
                     "state_l2": float(np.sqrt(sum(np.vdot(a, a).real for a in u))),
                     "spectrum": spec.tolist()})
        print(f"step {chk}: t={t:.9f} dt={dt:.6e} Ekin={diag[0]:.12e} ({time.time() - t0:.0f} s)", flush=True)
    doc = {"workload": "128^3 FP64 forced box, nu=0.01, f=(sin 4y,0,0), CFL 0.5 controller, seed 0, k0=4, E=0.5",
           "n": N, "nu": NU, "dt0": DT0, "cfl_min": CFL, "cfl_max": CFL, "dt_max": DTMAX, "kolmo_amp": 1.0,
           "kolmo_k": 4, "seed": 0, "k0": 4.0, "energy": 0.5, "steps": STEPS, "every": EVERY,
           "generator": "oracle/dnsbox_oracle.c orc_heun_adaptive + orc_diagnostics", "checkpoints": rows}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
