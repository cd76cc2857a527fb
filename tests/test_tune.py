"""Host-side checks of the autotuner: the table round-trips, candidates respect the kernels'
constraints, and the LDS bank model ranks the padded x-pass layout above the plain one."""
import os
import tempfile

from dnsbox_amd.tune import autotune, cfgtable, ldsmodel


def prod(p):
    out = 1
    for r in p:
        out *= r
    return out


def test_table_round_trip():
    t = cfgtable.parse()
    assert set(t) == {(n, p) for n in cfgtable.SIZES for p in ("float", "double")}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.hpp")
        cfgtable.emit(t, ["note"], path)
        assert cfgtable.parse(path) == t


def test_table_entries_are_consistent():
    for (n, prec), c in cfgtable.parse().items():
        assert prod(c.zp) == n and prod(c.yp) == n and prod(c.rp) == n // 2
        assert c.zp[0] == c.zp[-1], "z plans are symmetric"
        assert (n * n) % c.rb == 0
        if n in cfgtable.GROUP_SIZES:
            assert c.zw * n // max(c.zp) == 256 and c.yw * n // max(c.yp) == 256 and c.ntx == 256


def test_candidates_respect_constraints():
    t = cfgtable.parse()
    for n, prec in ((32, 32), (48, 32), (64, 64), (256, 32)):
        esz = 8 if prec == 32 else 16
        base = t[(n, "float" if prec == 32 else "double")]
        for c in autotune.z_candidates(n, esz, base):
            assert prod(c.zp) == n and c.zp[0] == c.zp[-1]
            assert all(max(c.zp) % r == 0 for r in c.zp)
            assert autotune._threads_ok(n, c.zw * n // max(c.zp))
        for c in autotune.y_candidates(n, esz, base):
            assert prod(c.yp) == n and autotune._threads_ok(n, c.yw * n // max(c.yp))
        if n <= 48:   # the bank model is slow at large n
            for c in autotune.x_candidates(n, prec, base, limit=3):
                assert prod(c.rp) == n // 2 and (n * n) % c.rb == 0 and c.ntx == 256


def test_bank_model_prefers_padding_at_256():
    plain = ldsmodel.xpass_ratio(256, 8, 4, 256, (16, 8), 129, 0)
    padded = ldsmodel.xpass_ratio(256, 8, 4, 256, (16, 8), 136, 32)
    assert plain > 1.9 and padded < 1.3


def test_choose_combines_axes_and_rejects_wrong_energy():
    base = {"zw": 16, "yw": 64, "rb": 4, "ntx": 256, "zp": [8, 4, 8], "yp": [16, 16], "rp": [16, 8],
            "xpd": 32, "xpad": 2}
    t0 = {"zinv_curl": 1, "yinv6": 1, "xcross": 1, "yfwd3": 1, "zpred_zinv": 1, "zcorr_zinv": 1, "step": 9,
          "ekin": 0.5}
    mk = lambda axis, i, cfg, **kw: {"n": 256, "prec": 32, "state": 64, "id": i, "axis": axis, "cfg": cfg,
                                     "t": dict(t0, **kw)}
    rs = [mk("base", 0, base),
          mk("z", 1, dict(base, zw=8), zpred_zinv=0.5),
          mk("y", 2, dict(base, yw=32), yinv6=0.5),
          mk("x", 3, dict(base, rb=8), xcross=0.1, ekin=0.6)]   # wrong energy: rejected
    cfg, notes = autotune.choose(rs)[(256, 32, 64)]
    assert cfg["zw"] == 8 and cfg["yw"] == 32 and cfg["rb"] == 4
    assert any("rejected" in n for n in notes)
