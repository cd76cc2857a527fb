"""dnsbox_rhs_dev_noalloc (no context, no device allocation; its scratch is a static device array)
against dnsbox_rhs_dev of a context with the same parameters, through the C-ABI: the same kernels
on the same inputs, so the same bits.  Two calls in a row, the second on a smaller box, so the
second call runs on the array the first one left behind."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import Oracle, to_torch_layout

pytestmark = pytest.mark.gpu


def dev(f, dtype):
    import torch
    return torch.from_numpy(np.stack([to_torch_layout(a) for a in f])).to(dtype).cuda().contiguous()


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_rhs_matches_context_rhs(gpu_lib, prec):
    import torch
    from dnsbox_amd import Box
    from dnsbox_amd._lib import check
    from dnsbox_amd.box import default_params
    dtype = torch.complex128 if prec == 64 else torch.complex64
    cases = [(64, dict(nu=0.02, kolmo_amp=1.0, kolmo_k=4, bandlim=0)),
             (32, dict(nu=0.01, Pin=0.1, kcut=2.5, bandlim=1, alpha_z=0.5))]
    for n, kw in cases:
        o = Oracle(n)
        f = o.init_solenoidal(5, 0, 3.0, 0.5)
        u = dev(f, dtype)
        want = torch.empty_like(u)
        with Box(n=n, precision=prec, adaptive_dt=0, **kw) as b:
            b.rhs_dev(u, want)
            torch.cuda.synchronize()
        p = default_params(n=n, precision=prec, adaptive_dt=0, **kw)
        got = torch.full_like(u, float("nan"))
        stream = torch.cuda.current_stream().cuda_stream
        print(f"noalloc n={n} prec={prec} {kw}", flush=True)
        check(gpu_lib.dnsbox_rhs_dev_noalloc(C.byref(p), C.c_void_p(u.data_ptr()), C.c_void_p(got.data_ptr()),
                                             C.c_void_p(stream)))
        torch.cuda.synchronize()
        diff = (got - want).abs().max().item()
        print(f"  max |noalloc - context| = {diff:.3e}, max |want| = {want.abs().max().item():.3e}", flush=True)
        assert torch.equal(got, want)


def test_noalloc_rejects_large_boxes(gpu_lib):
    from dnsbox_amd.box import default_params
    p = default_params(n=128, precision=64, adaptive_dt=0)
    rc = gpu_lib.dnsbox_rhs_dev_noalloc(C.byref(p), C.c_void_p(1), C.c_void_p(1), None)
    assert rc == -1 and b"n must be 16, 32, 48 or 64" in gpu_lib.dnsbox_last_error()
