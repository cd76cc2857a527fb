"""Run control around the GPU step: the time loop of dnsboxheun.f90 with its files.

    run(box, outdir, nt, ...)

reproduces the driver around the hot path (dnsboxheun.f90:18-189):
  * info.dat (io_saveInfo, io.f90:704-735), Ekin.dat rows (io_save_stats, io.f90:628-661:
    time, Ekin, Disp, 2 Q Ekin, Courant, divMax, EZero, Eband as 8e20.12), specNNNN.dat
    (io_save_spectrum, io.f90:680-702: kSpec, Espec as 2e20.12, appended), in the
    reference's formats so that python/analysis.py reads a run made here;
  * the state at t = 0 and every iSaveRate1 steps as stateNNNN.h5 (io_save_state,
    io.f90:100-278: datasets u, v, w, the physical velocity, FP64, Fortran (Nx, Ny, Nz)), through
    dnsbox_amd.state_io; a run can start from such a file (initial_state, as io_load_state at
    dnsboxheun.f90:48);
  * the RUNNING sentinel: created at the start, the run stops cleanly when it is deleted;
  * the step-size policy: with tstep_fix the run aborts when the Courant number exceeds
    courant_max (saving the final spectrum and state); otherwise set_time_step
    (dnsboxheun.f90:227-246) rescales dt to Courant (cmin+cmax)/2, capped at tstep_max, when
    the Courant number of the last statistics leaves [cmin, cmax] — policy="reference".
    policy="device" uses the box's own per-step controller instead (adaptive_dt = 1).
Steps between two events (statistics, saves, the end) run as one dnsbox_step call, so the
sentinel is looked at every iSaveRate2 steps (every step in the reference).
"""
from __future__ import annotations

import math
import os
import time as _time

import numpy as np

from . import _lib


def fortran_e(x, w=20, d=12):
    """Fortran Ew.d edit descriptor: 0.ddddE+xx right-aligned in w columns."""
    x = float(x)
    if x == 0.0 or not math.isfinite(x):
        body = f"0.{'0' * d}E+00" if x == 0.0 else str(x)
        return body.rjust(w)
    e = math.floor(math.log10(abs(x))) + 1
    m = abs(x) / 10.0 ** e
    digits = round(m * 10 ** d)
    if digits >= 10 ** d:          # rounding carried into a new digit
        digits //= 10
        e += 1
    sign = "-" if x < 0 else ""
    ex = f"{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) < 100 else f"{e:+d}"
    return f"{sign}0.{digits:0{d}d}E{ex}".rjust(w)


def fortran_f(x, w=12, d=9):
    return f"{float(x):{w}.{d}f}"


def write_info(path, box, isave1, isave2, dt, courant, initrand=True):
    """io_saveInfo (io.f90:704-735), appended."""
    p = box.params
    lines = [f"Nx,y,z = {box.nx:4d}{box.ny:4d}{box.nz:4d}",
             f"iSaveRate1 = {isave1:6d}",
             f"iSaveRate2 = {isave2:6d}",
             "alpha_x,y,z = " + "".join(fortran_f(float(np.float32(a or 1.0)))
                                        for a in (getattr(p, "alpha_x", 1.0), getattr(p, "alpha_y", 1.0),
                                                  getattr(p, "alpha_z", 1.0))),
             "dt = " + fortran_f(dt),
             "nu = " + fortran_f(p.nu)]
    if p.bandlim:
        lines += ["Pin = " + fortran_f(p.Pin), "kF = " + fortran_f(p.kcut)]
    else:
        lines += ["Q = " + fortran_f(p.Q)]
    lines += ["Deltak = " + fortran_f(p.deltak or 1.0), "Courant = " + fortran_f(courant),
              " Random initial condition" if initrand else " Initial condition read from file"]
    with open(path, "a") as f:
        f.write("\n".join(lines) + "\n")


class Run:
    """State of one run directory (counters and open files of io.f90)."""

    def __init__(self, box, outdir):
        self.box = box
        self.dir = outdir
        os.makedirs(outdir, exist_ok=True)
        self.save_count1 = 0
        self.courant = 0.0
        self.ekin = open(os.path.join(outdir, "Ekin.dat"), "a")

    def path(self, name):
        return os.path.join(self.dir, name)

    def save_stats(self):
        d, _ = self.box.diagnostics()
        self.courant = float(d[_lib.D_COURANT])
        row = (d[_lib.D_TIME], d[_lib.D_EKIN], d[_lib.D_DISP], d[_lib.D_PROD], d[_lib.D_COURANT],
               d[_lib.D_DIVMAX], d[_lib.D_EZERO], d[_lib.D_EBAND])
        self.ekin.write("".join(fortran_e(v) for v in row) + "\n")
        self.ekin.flush()
        return d

    def save_spectrum(self):
        _, espec = self.box.diagnostics()
        dk = self.box.params.deltak or 1.0
        with open(self.path(f"spec{self.save_count1:04d}.dat"), "a") as f:
            for nk, e in enumerate(espec, start=1):
                f.write(fortran_e(nk * dk) + fortran_e(e) + "\n")

    def save_state(self):
        self.box.save_state(self.path(f"state{self.save_count1:04d}.h5"))
        self.save_count1 += 1

    def close(self):
        self.ekin.close()


def set_time_step(box, courant, cmin, cmax, tstep_max):
    """dnsboxheun.f90:227-246: dt <- (cmin+cmax)/(2 C) dt, capped; returns the new dt."""
    _, dt = box.time()
    dt = ((cmax + cmin) / (2.0 * courant)) * dt
    dt = min(dt, tstep_max)
    box.rhs_int_fact(dt)
    return dt


def run(box, outdir, nt, isave1=1000, isave2=10, tstep_fix=True, courant_min=0.15, courant_max=0.2,
        tstep_max=0.01, policy="reference", running_file=True, initrand=True, log=None, initial_state=None):
    """The time loop of dnsboxheun.f90 on `box` (its state and dt already set, or the state read
    from the state file `initial_state`, then dealiased and projected as dnsboxheun.f90:48-55
    does); returns the number of steps taken and the reason the run stopped."""
    log = log or (lambda *a: None)
    if initial_state is not None:
        box.load_state(initial_state)
        initrand = False
    r = Run(box, outdir)
    _, dt = box.time()
    # dnsboxheun.f90:18-60: io_saveInfo before the first statistics (Courant not yet
    # computed), then spectrum, statistics and state of the initial field
    write_info(r.path("info.dat"), box, isave1, isave2, dt, 0.0, initrand)
    r.save_spectrum()
    r.save_stats()
    r.save_state()
    if not tstep_fix and policy == "reference":
        dt = set_time_step(box, r.courant, courant_min, courant_max, tstep_max)
    sentinel = r.path("RUNNING")
    if running_file:
        with open(sentinel, "w") as f:
            f.write(" This file indicates a job running in this directory.\n"
                    " Delete this file to cleanly terminate the process.\n")
    n = 0
    reason = "sentinel removed"
    t0 = _time.perf_counter()
    while (not running_file) or os.path.exists(sentinel):
        # steps up to the next event; one step when the (last) Courant number already
        # ends a fixed-dt run, which the reference checks after every step
        nxt = min(nt, (n // isave1 + 1) * isave1, (n // isave2 + 1) * isave2)
        if tstep_fix and r.courant > courant_max:
            nxt = n + 1
        box.step(nxt - n, 1)
        n = nxt
        if n % isave1 == 0:
            r.save_spectrum()
            r.save_state()
        if n % isave2 == 0:
            d = r.save_stats()
            log(f"n={n} t={d[_lib.D_TIME]:.6f} Ekin={d[_lib.D_EKIN]:.6e} Courant={r.courant:.4f}")
        if not tstep_fix and policy == "reference" and (
                r.courant > courant_max or (r.courant < courant_min and box.time()[1] < tstep_max)):
            dt = set_time_step(box, r.courant, courant_min, courant_max, tstep_max)
            r.courant = float(box.diagnostics()[0][_lib.D_COURANT])
            log(f"time step set to {dt:.6e}, new Courant number {r.courant:.4f}")
        elif tstep_fix and r.courant > courant_max:
            reason = "Courant number exceeded the maximum allowed value"
            r.save_spectrum()
            r.save_state()
            break
        elif n >= nt:
            reason = "maximum time steps reached"
            r.save_spectrum()
            r.save_state()
            break
    if running_file and os.path.exists(sentinel):
        os.remove(sentinel)
    r.close()
    log(f"{reason} after {n} steps ({n / max(_time.perf_counter() - t0, 1e-9):.1f} steps/s)")
    return n, reason
