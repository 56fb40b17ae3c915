"""GPU parity of the device-side layout build (HDDM_LB_OPT_DEVICE_LAYOUT = 1,
the default): every device array of an uploaded state — trie order, op
stream with its skips, surplus rows and splits, f16 column scales, support
masks, factor table — equals, byte for byte, the array uploaded from the host
build (option 0), and both give the same forward to the last bit.

Node sets: regular and adaptive grids in D = 3, 16, 20 and 59, dofs that are
and are not multiples of 32, nodes of 8 to 15 factors, and surpluses spanning
many orders of magnitude (f16 subnormals and rounding ties in the splits).
Indices the device build does not take (|i| >= 2^21) go to the host build,
which builds them or reports the error.
"""
import time

import numpy as np
import pytest

import hddm_lb
from test_gpu_deep import deep_node_set

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def upload_both(ii, jj, S):
    ctxs, ms = [], []
    for dev in (1, 0):
        ctx = hddm_lb.Context(0)
        ctx.set_option(hddm_lb.OPT_DEVICE_LAYOUT, dev)
        t = time.perf_counter()
        ctx.upload_state(0, ii, jj, S)
        ms.append((time.perf_counter() - t) * 1e3)
        ctxs.append(ctx)
    say("  upload %d nodes: device build %.1f ms, host build %.1f ms" % (ii.shape[0], ms[0], ms[1]))
    return ctxs


def check_same(torch, ii, jj, S, X):
    cd, ch = upload_both(ii, jj, S)
    for name in hddm_lb.STATE_ARRAYS:
        a, b = cd.state_array(0, name), ch.state_array(0, name)
        assert a.size == b.size, name
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            raise AssertionError("%s differs at %d of %d bytes, first at byte %d" % (name, bad.size, a.size, bad[0]))
    x = torch.from_numpy(X).cuda()
    for prec in (hddm_lb.FP64, hddm_lb.F32, hddm_lb.F16X2):
        ys = []
        for ctx in (cd, ch):
            y = torch.full((X.shape[0], S.shape[1]), np.nan, dtype=torch.float64, device="cuda")
            ctx.forward(x, y, None, prec)
            torch.cuda.synchronize()
            ys.append(y.cpu().numpy())
        assert np.array_equal(ys[0], ys[1]), prec
    cd.close()
    ch.close()


@pytest.mark.parametrize("dim,start,maxl,p,seed,target,dofs", [
    (3, 3, 4, 0.0, 1, 0, 8),
    (16, 2, 7, 0.3, 2, 3000, 32),
    (20, 1, 6, 0.3, 3, 5000, 40),
    (59, 3, 4, 0.0, 4, 0, 128),
])
def test_device_layout_matches_host(torch_mod, dim, start, maxl, p, seed, target, dofs):
    ii, jj, lv, _ = hddm_lb.generate_grid(dim, start, maxl, p, seed, target)
    S = hddm_lb.synthetic_surplus(lv, dofs, seed + 1)
    X = hddm_lb.synthetic_points(256, dim, seed + 2)
    say("layout D=%d nodes %d dofs %d" % (dim, ii.shape[0], dofs))
    check_same(torch_mod, ii, jj, S, X)


def test_device_layout_deep_nodes(torch_mod):
    ii, jj, lv = deep_node_set(21, 1500)
    S = hddm_lb.synthetic_surplus(lv, 24, 22)
    assert ((ii != 0) | (jj != 0)).sum(axis=1).max() >= 9
    X = hddm_lb.synthetic_points(128, 16, 23, 0.02, 0.45)
    check_same(torch_mod, ii, jj, S, X)


def test_device_layout_wide_range_surpluses(torch_mod):
    """Surpluses from 1e-30 to 1e3 (with zeros, a zero column and exact f16
    halfway cases) exercise the f16 and bf16 splits' rounding."""
    ii, jj, lv, _ = hddm_lb.generate_grid(6, 2, 5, 0.4, 31, 1500)
    rng = np.random.default_rng(32)
    n = ii.shape[0]
    S = rng.standard_normal((n, 36)) * 10.0 ** rng.uniform(-30, 3, size=(n, 36))
    S[rng.random((n, 36)) < 0.05] = 0.0
    S[:, 5] = 0.0
    # values whose scaled f16 image lies exactly halfway between two f16 values
    S[:64, 7] = (1.0 + (2 * np.arange(64) + 1) * 2.0 ** -11) * 2.0 ** rng.integers(-8, 8, size=64)
    X = hddm_lb.synthetic_points(128, 6, 33)
    check_same(torch_mod, ii, jj, S, X)


def test_device_layout_defers_large_indices(torch_mod):
    """|i| = 2^22 is outside the device build's range: the upload falls back
    to the host build, which takes it (the FP64 value of 2^22 has a zero low
    word, all the walk reads) and gives the same arrays as option 0."""
    ii, jj, lv, _ = hddm_lb.generate_grid(4, 2, 4, 0.0, 41, 0)
    ii = ii.copy()
    jj = jj.copy()
    ii[-1, 0], jj[-1, 0] = 1 << 22, 1
    S = hddm_lb.synthetic_surplus(lv, 8, 42)
    X = hddm_lb.synthetic_points(64, 4, 43)
    check_same(torch_mod, ii, jj, S, X)


def test_device_layout_rejects_like_host(torch_mod):
    """i = 2^21 + 1 needs the FP64 low word: both builds refuse it with the
    same error, and the state slot stays empty."""
    ii, jj, lv, _ = hddm_lb.generate_grid(4, 2, 4, 0.0, 41, 0)
    ii = ii.copy()
    jj = jj.copy()
    ii[-1, 0], jj[-1, 0] = (1 << 21) + 1, 1
    S = hddm_lb.synthetic_surplus(lv, 8, 42)
    msgs = []
    for dev in (1, 0):
        ctx = hddm_lb.Context(0)
        ctx.set_option(hddm_lb.OPT_DEVICE_LAYOUT, dev)
        with pytest.raises(hddm_lb.HddmLbError) as e:
            ctx.upload_state(0, ii, jj, S)
        msgs.append(str(e.value))
        with pytest.raises(hddm_lb.HddmLbError):
            ctx.state_array(0, "ops")
        ctx.close()
    assert msgs[0] == msgs[1] and "2^21" in msgs[0]
