"""The allocation-free forward (hddm_lb_forward_noalloc) through the C ABI.

Its kernels are the forward's, run on scratch carved out of a static device
array instead of context allocations, so its output must equal
hddm_lb_forward's bit for bit on the same inputs: checked on two calls in a
row, the second with fewer points (the array keeps the first call's contents),
for a state split into node segments (partials in the array) and for several
states with invalid ids, plus the rejection of a batch too large for the
array.
"""
import ctypes

import numpy as np
import pytest

import hddm_lb

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def noalloc(ctx, x, y, sid, prec):
    rc = hddm_lb.lib.hddm_lb_forward_noalloc(ctx._h, x.shape[0], x.data_ptr(), x.stride(0),
                                             None if sid is None else sid.data_ptr(),
                                             y.data_ptr(), y.stride(0), prec, None)
    return rc


def pair(torch, ctx, X, sid, ldy, prec):
    x = torch.from_numpy(X).cuda()
    s = None if sid is None else torch.from_numpy(sid).cuda()
    y1 = torch.full((X.shape[0], ldy), np.nan, dtype=torch.float64, device="cuda")
    y2 = torch.full((X.shape[0], ldy), np.nan, dtype=torch.float64, device="cuda")
    ctx.forward(x, y1, s, prec)
    rc = noalloc(ctx, x, y2, s, prec)
    assert rc == 0, hddm_lb.last_error()
    torch.cuda.synchronize()
    return y1.cpu().numpy(), y2.cpu().numpy()


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32, hddm_lb.F16X2])
def test_noalloc_matches_forward_two_calls(torch_mod, prec):
    """A state split into node segments (partials live in the array): the
    second call, with fewer points, sees the first call's leftovers."""
    ii, jj, lv, _ = hddm_lb.generate_grid(12, 2, 5, 0.4, 51, 20000)
    S = hddm_lb.synthetic_surplus(lv, 40, 52)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    for n in (3000, 1100):
        X = hddm_lb.synthetic_points(n, 12, 53 + n)
        a, b = pair(torch_mod, ctx, X, None, 40, prec)
        say("noalloc n", n, "precision", prec, "max |diff|", np.nanmax(np.abs(a - b)))
        assert np.isfinite(b).all()
        assert np.array_equal(a, b)
    ctx.close()


def test_noalloc_multistate_invalid_ids(torch_mod):
    specs = [(10, 2, 4, 0.5, 61, 800, 16), (10, 2, 5, 0.4, 62, 3000, 64)]
    ctx = hddm_lb.Context(0)
    for s, (d, st, ml, p, seed, tgt, nd) in enumerate(specs):
        ii, jj, lv, _ = hddm_lb.generate_grid(d, st, ml, p, seed, tgt)
        ctx.upload_state(s, ii, jj, hddm_lb.synthetic_surplus(lv, nd, seed + 1))
    for n in (2000, 700):
        X = hddm_lb.synthetic_points(n, 10, 63 + n)
        sid = np.random.default_rng(n).integers(-1, 3, size=n).astype(np.int32)
        a, b = pair(torch_mod, ctx, X, sid, 70, hddm_lb.F32)
        assert np.array_equal(a, b, equal_nan=True)
        assert np.isnan(b[(sid < 0) | (sid > 1)]).all()
    ctx.close()


def test_noalloc_rejects_oversize_batch(torch_mod):
    torch = torch_mod
    ii, jj, lv, _ = hddm_lb.generate_grid(6, 2, 4, 0.5, 71, 0)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, hddm_lb.synthetic_surplus(lv, 8, 72))
    n = 1 << 24   # 8 bytes of scratch per point: 128 MiB, above the 64 MiB array
    x = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
    y = torch.zeros((n, 8), dtype=torch.float64, device="cuda")
    rc = noalloc(ctx, x, y, None, hddm_lb.FP64)
    assert rc == 1 and "scratch" in hddm_lb.last_error()
    ctx.close()
