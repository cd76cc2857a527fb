"""GPU FP64 path vs the CPU oracle (oracle/dnsbox_oracle.c), through the C-ABI.

Tolerances (FP64): RHS <= 1e-12 relative L2; state after 100 IF-Heun steps
<= 1e-9 relative L2; spectrum shell by shell <= 1e-10 relative.
"""
import numpy as np
import pytest

from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = sum(np.linalg.norm((x - y).ravel()) ** 2 for x, y in zip(a, b))
    den = sum(np.linalg.norm(y.ravel()) ** 2 for y in b)
    return float(np.sqrt(num / den))


def make(n, **kw):
    from dnsbox_amd import Box
    return Box(n=n, precision=kw.pop("precision", 64), **kw)


@pytest.mark.parametrize("n", [16, 32, 48])
def test_init_random_matches_oracle(gpu_lib, n):
    o = Oracle(n)
    ref = o.init_solenoidal(7, 0, 4.0, 0.5)
    with make(n, adaptive_dt=0, bandlim=0) as b:
        b.init_random(7, 0, 4.0, 0.5)
        got = b.get_state()
    print(f"init n={n} rel_l2={rel_l2(got, ref):.3e}", flush=True)
    assert rel_l2(got, ref) < 1e-13


@pytest.mark.parametrize("n", [16, 32, 48, 64])
def test_rhs_matches_oracle(gpu_lib, n):
    kw = dict(nu=0.05, kolmo_amp=1.0, kolmo_k=4 if n >= 16 else 1, bandlim=0)
    o = Oracle(n, nu=0.05, kolmo_amp=1.0, kolmo_k=kw["kolmo_k"])
    f = o.init_solenoidal(3, 0, 4.0, 0.5)
    ref = o.rhs(f)
    with make(n, adaptive_dt=0, **kw) as b:
        b.set_state(*f)
        got = b.rhs_nonlinear()
    err = rel_l2(got, ref)
    print(f"rhs n={n} rel_l2={err:.3e}", flush=True)
    assert err <= 1e-12


def test_rhs_bandlimited_matches_oracle(gpu_lib):
    n = 32
    o = Oracle(n, nu=0.005, Pin=0.1, kcut=2.0, bandlim=True)
    f = o.init_solenoidal(5, 0, 3.0, 0.5)
    ref = o.rhs(f)
    with make(n, nu=0.005, Pin=0.1, kcut=2.0, bandlim=1, adaptive_dt=0) as b:
        b.set_state(*f)
        got = b.rhs_nonlinear()
    err = rel_l2(got, ref)
    print(f"rhs band rel_l2={err:.3e}", flush=True)
    assert err <= 1e-12


def test_heun_100_steps_matches_oracle(gpu_lib):
    """The 32^3 FP64 workload: nu = 0.05, Kolmogorov sin 4y, fixed dt = 0.005, 100 steps."""
    n = 32
    o = Oracle(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(0, 0, 4.0, 0.5)
    ref = o.heun(f, 100, 0.005)
    with make(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.005) as b:
        b.set_state(*f)
        b.rhs_int_fact(0.005)
        b.step(100)
        got = b.get_state()
        t, dt = b.time()
        gd, gs = b.diagnostics()
    err = rel_l2(got, ref)
    print(f"heun100 rel_l2={err:.3e} t={t}", flush=True)
    assert err <= 1e-9
    assert abs(t - 0.5) < 1e-12
    od, os_ = o.diagnostics(ref, 0.005)
    nz = os_ > 1e-30
    assert np.max(np.abs(gs[nz] - os_[nz]) / os_[nz]) < 1e-9
    for i in (0, 1, 2, 3, 5, 6, 7):
        assert abs(gd[i] - od[i]) <= 1e-9 * max(abs(od[i]), 1e-300) + 1e-13, (i, gd[i], od[i])


def test_diagnostics_match_oracle(gpu_lib):
    n = 32
    o = Oracle(n, nu=0.01, Q=0.1, Pin=0.1, kcut=2.5)
    f = o.init_solenoidal(11, 0, 4.0, 0.5)
    od, os_ = o.diagnostics(f, 0.004)
    with make(n, nu=0.01, Q=0.1, kcut=2.5, bandlim=0, adaptive_dt=0, dt=0.004) as b:
        b.set_state(*f)
        gd, gs = b.diagnostics()
    # Ekin, Disp, Prod, Courant, EZero, EOne, Esqrt2, Eband
    for i in (0, 1, 2, 3, 5, 6, 7, 8):
        assert abs(gd[i] - od[i]) <= 1e-11 * max(abs(od[i]), 1e-300) + 1e-14, (i, gd[i], od[i])
    assert abs(gd[4] - od[4]) < 1e-12  # divMax (rounding-level for a solenoidal field)
    np.testing.assert_allclose(gs, os_, rtol=1e-11, atol=1e-300)


def test_physical_roundtrip(gpu_lib):
    n = 32
    o = Oracle(n, nu=0.05)
    u = o.init_shapiro(0.0)
    with make(n, adaptive_dt=0, bandlim=0) as b:
        b.set_state_physical(*u)
        spec = b.get_state()
        back = b.get_state_physical()
    ref = [o.r2c(a) for a in u]
    assert rel_l2(spec, ref) < 1e-13
    assert max(np.abs(a - x).max() for a, x in zip(back, u)) < 1e-13


def test_shapiro_kat(gpu_lib):
    """state.f90:287-348: IF-Heun on the Shapiro/Beltrami field reproduces exp(-3 nu t)."""
    n = 32
    nu = 0.05
    o = Oracle(n, nu=nu)
    u = o.init_shapiro(0.0)
    with make(n, nu=nu, adaptive_dt=0, bandlim=0, dt=0.01) as b:
        b.set_state_physical(*u)
        b.state_dealias()
        b.state_project()
        b.rhs_int_fact(0.01)
        b.step(100)
        phys = b.get_state_physical()
    err = o.shapiro_error(1.0, phys)
    print(f"shapiro err={err:.3e}", flush=True)
    assert err < 1e-12


def test_fp32_rhs_close_to_fp64(gpu_lib):
    n = 64
    o = Oracle(n, nu=0.02, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(2, 0, 4.0, 0.5)
    ref = o.rhs(f)
    with make(n, precision=32, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0) as b:
        b.set_state(*f)
        got = b.rhs_nonlinear()
    err = rel_l2(got, ref)
    print(f"fp32 rhs rel_l2={err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("state_precision", [0, 64])
def test_fp32_256_matches_oracle(gpu_lib, state_precision):
    """The 256^3 FP32 step (the headline size; the step's passes run pruned) against the FP64
    oracle: the RHS, then two fixed-dt IF-Heun steps."""
    n, dt = 256, 0.002
    kw = dict(nu=0.005, kolmo_amp=1.0, kolmo_k=4)
    o = Oracle(n, **kw, nthreads=16)
    f = o.init_solenoidal(3, 0, 4.0, 0.5)
    ref = o.rhs(f)
    with make(n, precision=32, state_precision=state_precision, bandlim=0, adaptive_dt=0, dt=dt, **kw) as b:
        b.set_state(*f)
        got = b.rhs_nonlinear()
        err = rel_l2(got, ref)
        print(f"fp32 256 rhs rel_l2={err:.3e}", flush=True)
        assert err < 1e-5
        b.set_state(*f)
        b.rhs_int_fact(dt)
        b.step(2)
        got = b.get_state()
    ref = o.heun(f, 2, dt)
    err = rel_l2(got, ref)
    print(f"fp32 256 (state {state_precision or 32}) heun x2 rel_l2={err:.3e}", flush=True)
    assert err < 1e-6


def test_fp32_512_matches_fp64_box(gpu_lib):
    """512^3: the FP32 transforms (register-resident x pass with 32 lanes per x-line, pruned
    passes) against the FP64 box (its x pass the LDS Stockham rows) on the same seeded field: the
    RHS, then two fixed-dt IF-Heun steps with the FP64 state.  The FP64 passes are pinned against
    the oracle at the sizes the oracle finishes in seconds."""
    n, dt = 512, 0.001
    kw = dict(nu=0.002, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt)
    with make(n, precision=64, **kw) as b:
        b.init_random(3, 0, 4.0, 0.5)
        f = b.get_state()
        ref_rhs = b.rhs_nonlinear()
        b.rhs_int_fact(dt)
        b.step(2)
        ref = b.get_state()
    with make(n, precision=32, state_precision=64, **kw) as b:
        b.set_state(*f)
        got = b.rhs_nonlinear()
        err = rel_l2(got, ref_rhs)
        print(f"fp32 512 rhs vs fp64 box rel_l2={err:.3e}", flush=True)
        assert err < 1e-5
        b.set_state(*f)
        b.rhs_int_fact(dt)
        b.step(2)
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"fp32 512 (state 64) heun x2 vs fp64 box rel_l2={err:.3e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("n,prec,sp,band", [(16, 64, 0, 0), (32, 64, 0, 0), (32, 64, 0, 1), (48, 32, 0, 0),
                                             (64, 64, 0, 0), (128, 32, 64, 0), (256, 32, 64, 0), (256, 64, 0, 0),
                                             (512, 32, 64, 0)])
def test_pruned_step_matches_unpruned(gpu_lib, n, prec, sp, band):
    """The step's passes pruned to the 2/3-rule cube skip only exact zeros: the results agree
    with the unpruned passes to rounding (state, clock, adaptive dt, diagnostics).  Not to the
    bit: the two kernel instantiations may contract multiply-adds differently."""
    kw = dict(nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=band, Pin=0.1, kcut=2.0, adaptive_dt=1,
              cfl_min=0.5, cfl_max=0.5, dt=0.002)
    tol = 1e-13 if prec == 64 else (1e-8 if sp == 64 else 1e-6)
    with make(n, precision=prec, state_precision=sp, **kw) as a, \
            make(n, precision=prec, state_precision=sp, **kw) as b:
        b.set_pruning(False)
        for x in (a, b):
            x.init_random(5, 0, 4.0, 0.5)
        assert a.pruning() and not b.pruning()
        for x in (a, b):
            x.step(12)
        sa, sb = a.get_state(), b.get_state()
        err = rel_l2(sa, sb)
        (da, ea), (db, eb) = a.diagnostics(), b.diagnostics()
        derr = float(np.linalg.norm(ea - eb) / np.linalg.norm(eb))
        (ta, dta), (tb, dtb) = a.time(), b.time()
        print(f"pruned n={n} prec={prec} state={sp or prec} band={band}: state rel_l2={err:.3e} "
              f"spectrum rel_l2={derr:.3e} t {ta!r} vs {tb!r}", flush=True)
        assert err <= tol
        assert derr <= tol
        tt = 1e-12 if prec == 64 else 1e-6    # the CFL step follows max|u| of the transform type
        assert abs(ta - tb) <= tt * tb and abs(dta - dtb) <= tt * dtb


def test_pruning_follows_the_state(gpu_lib):
    """A field with energy outside the kept cube turns pruning off (the step then matches the
    oracle, masked modes included); state_dealias turns it back on."""
    n, dt = 32, 0.005
    o = Oracle(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(0, 0, 4.0, 0.5)
    g = [x.copy() for x in f]
    g[0][12, 3, 5] = 1e-3 + 2e-3j   # kx = 12 > (32 - 1) // 3: outside the kept cube
    with make(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=dt) as b:
        b.set_state(*f)
        assert b.pruning()
        b.set_state(*g)
        assert not b.pruning()
        b.rhs_int_fact(dt)
        b.step(3)
        err = rel_l2(b.get_state(), o.heun(g, 3, dt))
        print(f"unpruned step with masked-mode energy: rel_l2={err:.3e}", flush=True)
        assert err <= 1e-12
        b.state_dealias()
        assert b.pruning()
        b.state_ptr()
        assert not b.pruning()


def test_band_forcing_after_unpruned_steps(gpu_lib):
    """Band forcing with kCutOff beyond the kept cube (kcut = 6.5 > (N-1)/3 = 5 at 16^3): unpruned
    steps of a state outside the cube leave values outside the cube in the predicted-state
    buffer; after a dealias the pruned steps must not count them in the band energy of u*.
    Compared with the same run with pruning off throughout."""
    n, dt = 16, 0.004
    kw = dict(nu=0.02, Pin=0.1, kcut=6.5, bandlim=1, adaptive_dt=0, dt=dt)
    o = Oracle(n, nu=0.02)
    f = o.init_solenoidal(3, 0, 3.0, 0.5)
    rng = np.random.default_rng(2)
    f[1][6, 2, 1] = 0.05 * np.abs(f[1]).max() * (rng.standard_normal() + 1j)   # kx = 6: outside
    with make(n, **kw) as a, make(n, **kw) as b:
        b.set_pruning(False)
        for x in (a, b):
            x.set_state(*f)
        assert not a.pruning()
        for x in (a, b):
            x.step(5)
            x.state_dealias()
        assert a.pruning() and not b.pruning()
        for x in (a, b):
            x.step(10)
        err = rel_l2(a.get_state(), b.get_state())
    print(f"band after unpruned phase: pruned vs unpruned rel_l2={err:.2e}", flush=True)
    assert err < 1e-13


def test_adaptive_heun_matches_oracle(gpu_lib):
    """CFL controller (dnsboxheun.f90:153-157, 227-246) evaluated on the current field."""
    n = 32
    o = Oracle(n, nu=0.01, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(1, 0, 4.0, 0.5)
    ref, dt_ref, t_ref = o.heun_adaptive(f, 40, 0.003, 0.5, 0.5, 0.01)
    with make(n, nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=1, cfl_min=0.5,
              cfl_max=0.5, dt_max=0.01, dt=0.003) as b:
        b.set_state(*f)
        b.step(40)
        got = b.get_state()
        t, dt = b.time()
    print(f"adaptive: t={t} ({t_ref}) dt={dt} ({dt_ref}) err={rel_l2(got, ref):.3e}", flush=True)
    assert abs(t - t_ref) < 1e-12 and abs(dt - dt_ref) < 1e-13
    assert rel_l2(got, ref) < 1e-10


def test_bandlimited_heun_matches_oracle(gpu_lib):
    n = 32
    o = Oracle(n, nu=0.005, Pin=0.1, kcut=2.0, bandlim=True)
    f = o.init_solenoidal(6, 0, 3.0, 0.5)
    ref = o.heun(f, 50, 0.004)
    with make(n, nu=0.005, Pin=0.1, kcut=2.0, bandlim=1, adaptive_dt=0, dt=0.004) as b:
        b.set_state(*f)
        b.step(50)
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"band heun rel_l2={err:.3e}", flush=True)
    assert err < 1e-10


def test_ab2_matches_oracle(gpu_lib):
    n = 32
    o = Oracle(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4)
    f = o.init_solenoidal(4, 0, 4.0, 0.5)
    ref = o.ab2(f, 60, 0.005)
    with make(n, nu=0.05, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.005) as b:
        b.set_state(*f)
        b.step(60, stepper=2)
        got = b.get_state()
    err = rel_l2(got, ref)
    print(f"ab2 rel_l2={err:.3e}", flush=True)
    assert err < 1e-10


def test_step_pipeline_is_call_invariant(gpu_lib):
    """The fused z pass keeps the next RHS's z-inverse in scratch across dnsbox_step calls;
    splitting a run into calls, or interleaving diagnostics, changes nothing."""
    n = 32
    kw = dict(nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=1, cfl_min=0.5,
              cfl_max=0.5, dt=0.003)
    with make(n, **kw) as a, make(n, **kw) as b:
        a.init_random(9, 0, 4.0, 0.5)
        b.init_random(9, 0, 4.0, 0.5)
        a.step(50)
        b.step(20)
        b.diagnostics()
        b.step(13)
        b.rhs_nonlinear()
        b.step(17)
        sa, sb = a.get_state(), b.get_state()
        assert all(np.array_equal(x, y) for x, y in zip(sa, sb))
        assert a.time() == b.time()


def test_fp32_heun_close_to_fp64(gpu_lib):
    n = 64
    kw = dict(nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.004)
    with make(n, precision=64, **kw) as a, make(n, precision=32, **kw) as b:
        for x in (a, b):
            x.init_random(2, 0, 4.0, 0.5)
            x.step(100)
        (d64, s64), (d32, s32) = a.diagnostics(), b.diagnostics()
    rel = np.abs(s32 - s64)[s64 > 0] / s64[s64 > 0]
    print(f"fp32 vs fp64 after 100 steps: energy {abs(d32[0]-d64[0])/d64[0]:.2e} "
          f"spectrum max {rel.max():.2e}", flush=True)
    assert abs(d32[0] - d64[0]) / d64[0] < 1e-5
    assert rel[:20].max() < 1e-4


def test_fp32_state64_closer_to_fp64(gpu_lib):
    """FP32 transforms with the state kept in FP64 track the FP64 run more closely than an
    FP32 state does."""
    n = 64
    kw = dict(nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=0, dt=0.004)
    res = {}
    for name, prec, sp in (("f64", 64, 0), ("f32", 32, 0), ("f32s64", 32, 64)):
        with make(n, precision=prec, state_precision=sp, **kw) as b:
            b.init_random(2, 0, 4.0, 0.5)
            b.step(200)
            res[name] = b.get_state()
    e32 = rel_l2(res["f32"], res["f64"])
    e64s = rel_l2(res["f32s64"], res["f64"])
    print(f"state error vs fp64 after 200 steps: fp32 state {e32:.2e}, fp64 state {e64s:.2e}", flush=True)
    assert e64s < e32
    assert e64s < 1e-5


def test_time_passes_leaves_box_unchanged(gpu_lib):
    """The in-step pass timer works on a copy: state, time and dt are untouched, and every
    pass the step launches gets a positive duration (at 32^3 the y-inverse, x and y-forward
    passes are one plane launch, reported as the x pass, with no bytes of their own)."""
    from dnsbox_amd import Box
    with Box(n=32, precision=64, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=1,
             cfl_min=0.5, cfl_max=0.5, dt_max=0.01, dt=0.004) as b:
        b.init_random(5, 0, 4.0, 0.5)
        b.step(3)
        before = b.get_state()
        t0 = b.time()
        ms = b.time_passes(4)
        nbytes = b.pass_bytes()
        after = b.get_state()
        assert b.time() == t0
        assert all(np.array_equal(x, y) for x, y in zip(before, after))
        assert all(ms[k] > 0 for k, v in nbytes.items() if v > 0), (ms, nbytes)
        assert all(ms[k] == 0 for k, v in nbytes.items() if v == 0), (ms, nbytes)
        b.step(2)
        with Box(n=32, precision=64, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=1,
                 cfl_min=0.5, cfl_max=0.5, dt_max=0.01, dt=0.004) as ref:
            ref.init_random(5, 0, 4.0, 0.5)
            ref.step(5)
            assert all(np.array_equal(x, y) for x, y in zip(b.get_state(), ref.get_state()))


@pytest.mark.parametrize("prec,sp", [(64, 0), (32, 64)])
def test_graph_replay_matches_eager_steps(gpu_lib, prec, sp):
    """The hipGraph-replayed step equals the eager step bit for bit, over several replays
    with eager work (diagnostics, a state round trip) and synchronisations in between."""
    import os
    from dnsbox_amd import Box
    kw = dict(n=32, precision=prec, state_precision=sp, nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0,
              adaptive_dt=1, cfl_min=0.5, cfl_max=0.5, dt_max=0.01, dt=0.004)
    eager = Box(**kw)
    eager.init_random(9, 0, 3.0, 0.5)
    eager.step(1)
    os.environ["DNSBOX_STEP_GRAPH"] = "1"      # read at the box's first step
    try:
        graph = Box(**kw)
        graph.init_random(9, 0, 3.0, 0.5)
        graph.step(1)
    finally:
        del os.environ["DNSBOX_STEP_GRAPH"]
    for chunk in (16, 3, 40, 17):   # 16 steps per graph: replays, remainders, eager chunks
        eager.step(chunk)
        graph.step(chunk)
        d_e, s_e = eager.diagnostics()
        d_g, s_g = graph.diagnostics()
        # the shell sums have a fixed order: the same bits
        assert np.array_equal(d_e, d_g) and np.array_equal(s_e, s_g)
        a, b = eager.get_state(), graph.get_state()
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), chunk
        assert eager.time() == graph.time()
    graph.set_state(*graph.get_state())     # restarts the pipeline: the next step runs eagerly first
    eager.set_state(*eager.get_state())
    eager.step(33)
    graph.step(33)
    assert all(np.array_equal(x, y) for x, y in zip(eager.get_state(), graph.get_state()))
    # a change of pass mode between replays re-captures the graph: unpruned steps, a state with
    # energy outside the kept cube, then pruned steps again after a dealias
    rng = np.random.default_rng(1)
    for phase in ("unpruned", "outside", "pruned"):
        if phase == "unpruned":
            for x in (eager, graph):
                x.set_pruning(False)
        elif phase == "outside":
            s0 = eager.get_state()
            s0[0][12, 3, 5] += 1e-3 * (1 + 1j) * np.abs(s0[0]).max() * rng.standard_normal()   # kx = 12 > 10
            for x in (eager, graph):
                x.set_pruning(True)
                x.set_state(*s0)
            assert not graph.pruning()
        else:
            for x in (eager, graph):
                x.state_dealias()
            assert graph.pruning()
        eager.step(33)
        graph.step(33)
        assert all(np.array_equal(x, y) for x, y in zip(eager.get_state(), graph.get_state())), phase
        assert eager.time() == graph.time()
    eager.close()
    graph.close()


def test_128_fp64_adaptive_forced_box_matches_oracle(gpu_lib):
    """128^3 FP64 forced box, nu = 0.01, f = (sin 4y, 0, 0), CFL 0.5 controller: state, time, dt,
    energy, dissipation, production, Courant number and the shell spectrum against the oracle at
    checkpoints every 10 steps (20 steps: the CPU oracle takes ~1 s per step at this size)."""
    n = 128
    o = Oracle(n, nu=0.01, kolmo_amp=1.0, kolmo_k=4, nthreads=16)
    ref = o.init_solenoidal(0, 0, 4.0, 0.5)
    dt_ref, t_ref = 0.002, 0.0
    with make(n, nu=0.01, kolmo_amp=1.0, kolmo_k=4, bandlim=0, adaptive_dt=1, cfl_min=0.5,
              cfl_max=0.5, dt_max=0.01, dt=0.002) as b:
        b.set_state(*ref)
        for chk in (10, 20):
            ref, dt_ref, t_ref = o.heun_adaptive(ref, 10, dt_ref, 0.5, 0.5, 0.01, t_ref)
            b.step(10)
            got = b.get_state()
            t, dt = b.time()
            err = rel_l2(got, ref)
            od, os_ = o.diagnostics(ref, dt_ref)
            gd, gs = b.diagnostics()
            # shells past the dealiasing cutoff hold only rounding-level energy: compare them
            # against the largest shell
            floor = 1e-13 * os_.max()
            shell = np.max(np.abs(gs - os_) / np.maximum(os_, floor))
            print(f"128^3 step {chk}: rel_l2={err:.2e} t={t:.12f} ({t_ref:.12f}) dt={dt:.3e} "
                  f"Ekin={gd[0]:.12e} ({od[0]:.12e}) max shell err={shell:.2e}", flush=True)
            assert err < 1e-10
            assert abs(t - t_ref) < 1e-12 and abs(dt - dt_ref) < 1e-13
            for i in (0, 1, 2, 3):   # Ekin, Disp, Prod, Courant
                assert abs(gd[i] - od[i]) <= 1e-10 * max(abs(od[i]), 1e-300) + 1e-14, (i, gd[i], od[i])
            np.testing.assert_allclose(gs, os_, rtol=1e-10, atol=floor)


def test_band_stepper_with_reprojection_matches_oracle(gpu_lib):
    """stepper 3 (dnsboxband.f90:91-156): band-forced IF-Heun step, then state_project and
    state_dealias of the new state, every step."""
    n, dt, steps = 32, 0.004, 30
    o = Oracle(n, nu=0.005, Pin=0.1, kcut=2.0, bandlim=True)
    ref = o.init_solenoidal(12, 0, 3.0, 0.5)
    for _ in range(steps):
        ref = o.dealias(o.project(o.heun(ref, 1, dt)))
    with make(n, nu=0.005, Pin=0.1, kcut=2.0, bandlim=1, adaptive_dt=0, dt=dt) as b:
        b.set_state(*o.init_solenoidal(12, 0, 3.0, 0.5))
        b.step(steps, stepper=3)
        got = b.get_state()
        t, _ = b.time()
    err = rel_l2(got, ref)
    print(f"band stepper with reprojection rel_l2={err:.3e}", flush=True)
    assert err < 1e-10 and abs(t - steps * dt) < 1e-12
