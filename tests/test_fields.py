"""Initial fields and the Shapiro known-answer test (dnsbox_amd/fields.py)."""
import numpy as np
import pytest

from oracle_lib import Oracle


def test_shapiro_matches_oracle_and_is_solenoidal():
    from dnsbox_amd import fields
    n, nu, t = 16, 0.05, 0.3
    o = Oracle(n, nu=nu)
    for a, b in zip(fields.shapiro(n, nu, t), o.init_shapiro(t)):
        assert np.array_equal(a, b) or np.max(np.abs(a - b)) < 1e-15
    u, v, w = (np.fft.rfftn(a) for a in fields.shapiro(n, nu, 0.0))
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = k.reshape(-1, 1, 1), k.reshape(1, -1, 1), k[: n // 2 + 1].reshape(1, 1, -1)
    assert np.max(np.abs(kx * u + ky * v + kz * w)) < 1e-9 * n ** 3


def test_shell_index_rule():
    from dnsbox_amd.fields import shell_index
    absk = np.array([0.0, 0.49, 0.5, 1.0, 1.49, 1.5, 2.4999, np.sqrt(2.0), np.sqrt(3.0)])
    assert shell_index(absk).tolist() == [0, 0, 1, 1, 1, 2, 2, 1, 2]


@pytest.mark.gpu
def test_yoffe_init_has_the_target_spectrum(gpu_lib):
    from dnsbox_amd import Box, fields
    with Box(n=32, precision=64, nu=0.01, bandlim=0, adaptive_dt=0) as b:
        fields.init_yoffe(b, seed=3)
        d, espec = b.diagnostics()
        nk = np.arange(1, len(espec))
        target = fields.yoffe_spectrum(nk * 1.0)
        rel = np.max(np.abs(espec[:-1] - target) / target)
        print(f"yoffe spectrum rel err {rel:.2e}, divMax {d[4]:.2e}", flush=True)
        assert rel < 1e-12 and espec[-1] == 0.0
        assert abs(d[4]) < 1e-12
        a = b.get_state()
        fields.init_yoffe(b, seed=3, reference_w=False)
        assert not np.allclose(a[2], b.get_state()[2])


@pytest.mark.gpu
def test_shapiro_known_answer_through_fields(gpu_lib):
    from dnsbox_amd import Box, fields
    nu, dt, steps = 0.05, 0.01, 100
    with Box(n=32, precision=64, nu=nu, bandlim=0, adaptive_dt=0, dt=dt) as b:
        fields.init_shapiro(b, 0.0)
        assert fields.shapiro_error(b, 0.0) < 1e-14
        b.rhs_int_fact(dt)
        b.step(steps)
        err = fields.shapiro_error(b, steps * dt)
    print(f"shapiro KAT error {err:.2e}", flush=True)
    assert err < 1e-12
