"""FP32 transforms with the FP64 state against the FP64 step of the same field: energy spectrum
(shell by shell and L2) and energy after 1000 fixed-dt steps, at 256^3 (nu = 0.005) and 512^3
(nu = 0.002), the precision form of the 256^3 / 512^3 forced box (seeded E(k) ~ k^4 exp(-2 (k/4)^2),
E = 0.5, forcing sin 4y, dt about the CFL 0.5 step).  The FP64 run is the reference
(its own parity with the oracle is pinned in test_gpu_parity.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fixed_run(n, nu, prec, state_precision, steps, dt):
    from dnsbox_amd import Box
    with Box(n=n, precision=prec, state_precision=state_precision, nu=nu, kolmo_amp=1.0, kolmo_k=4,
             bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.init_random(0, 0, 4.0, 0.5)
        b.rhs_int_fact(dt)
        b.step(steps)
        return b.diagnostics()


def spectrum_errors(ref, got):
    (d64, s64), (d32, s32) = ref, got
    assert np.all(np.isfinite(s32)) and np.all(np.isfinite(s64))
    nz = s64 > 0
    shell = np.abs(s32[nz] - s64[nz]) / s64[nz]
    return (float(shell.max()), int(np.flatnonzero(nz)[shell.argmax()]),
            float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64)), float(abs(d32[0] - d64[0]) / d64[0]))


# n, nu, dt, bound on the worst shell, bound on the spectrum L2 and energy errors
CASES = [(256, 0.005, 0.002, 1e-5, 1e-6), (512, 0.002, 0.002 * 256 / 512 * 0.85, 1e-3, 1e-6)]


@pytest.mark.parametrize("n,nu,dt,shell_tol,tol", CASES)
def test_fp32_form_spectrum_vs_fp64_1000_steps(gpu_lib, n, nu, dt, shell_tol, tol):
    ref = fixed_run(n, nu, 64, 0, 1000, dt)
    got = fixed_run(n, nu, 32, 64, 1000, dt)
    worst, at, l2, en = spectrum_errors(ref, got)
    print(f"{n}^3 nu={nu} dt={dt:.3e}: worst shell {worst:.3e} (shell {at}), spectrum L2 {l2:.3e}, "
          f"energy {en:.3e}", flush=True)
    assert worst <= shell_tol and l2 <= tol and en <= tol
