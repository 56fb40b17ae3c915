"""GPU parity of the wave-specialized sweep (a walker wave feeding MFMA
consumer waves) against the oracle, forced with HDDM_LB_OPT_FORMULATION = 2.

Shapes cover one dof tile (two consumers splitting it), two to four tiles
(four consumers), more tiles than consumers (two passes), one, two and four
coordinate banks, the 16-row path stack, multistate with unequal dofs, node
segments, padded and invalid rows. After every call the device status word
must be clear: a wave that stopped waiting for another would have set it.
"""
import numpy as np
import pytest

import hddm_lb
import oracle
from gpu_util import ALL_PRECISIONS, TOL, normwise_rel_err, run_forward

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def ws_ctx():
    ctx = hddm_lb.Context(0)
    ctx.set_option(hddm_lb.OPT_FORMULATION, 2)
    ctx.device_status(clear=True)
    return ctx


@pytest.mark.parametrize("dim,dofs", [(8, 32), (8, 20), (20, 96), (33, 64), (59, 128), (16, 160)])
@pytest.mark.parametrize("prec", ALL_PRECISIONS)
def test_ws_forward_matches_oracle(torch_mod, dim, dofs, prec):
    ii, jj, lv, _ = hddm_lb.generate_grid(dim, 2, 5, 0.4, 40 + dim, 3000)
    S = hddm_lb.synthetic_surplus(lv, dofs, 41 + dim)
    X = hddm_lb.synthetic_points(333, dim, 42 + dim)
    ctx = ws_ctx()
    ctx.upload_state(0, ii, jj, S)
    y = run_forward(ctx, X, precision=prec)
    assert ctx.device_status() == 0
    yo = oracle.forward(ii, jj, S, X)
    err = normwise_rel_err(y, yo)
    say("ws dim", dim, "dofs", dofs, "nno", ii.shape[0], "precision", prec, "err", err)
    assert err <= TOL[prec]
    ctx.close()


@pytest.mark.parametrize("prec", ALL_PRECISIONS)
def test_ws_multistate_segments_padding(torch_mod, prec):
    """Three states of 16, 64 and 128 dofs (one to four tiles), forced node
    segments, a padded y (ldy 140) and invalid state ids: padding and
    invalid rows keep their NaN."""
    torch = torch_mod
    specs = [(24, 2, 5, 0.4, 70, 2500, 16), (24, 2, 5, 0.4, 71, 4000, 64), (24, 1, 5, 0.5, 72, 6000, 128)]
    ctx = ws_ctx()
    ctx.set_option(hddm_lb.OPT_SEGMENT_NODES, 700)
    states = []
    for s, (d, st, ml, p, seed, tgt, nd) in enumerate(specs):
        ii, jj, lv, _ = hddm_lb.generate_grid(d, st, ml, p, seed, tgt)
        S = hddm_lb.synthetic_surplus(lv, nd, seed + 1)
        ctx.upload_state(s, ii, jj, S)
        states.append((ii, jj, S))
    n = 900
    X = hddm_lb.synthetic_points(n, 24, 73)
    sid = np.random.default_rng(74).integers(-1, 4, size=n).astype(np.int32)  # -1 and 3 are invalid
    x = torch.from_numpy(X).cuda()
    sd = torch.from_numpy(sid).cuda()
    y = torch.full((n, 140), np.nan, dtype=torch.float64, device="cuda")
    ctx.forward(x, y, sd, prec)
    torch.cuda.synchronize()
    assert ctx.device_status() == 0
    Y = y.cpu().numpy()
    for s, (ii, jj, S) in enumerate(states):
        rows = sid == s
        err = normwise_rel_err(Y[rows, :S.shape[1]], oracle.forward(ii, jj, S, X[rows]))
        say("ws multistate state", s, "precision", prec, "err", err)
        assert err <= TOL[prec]
        assert np.isnan(Y[rows, S.shape[1]:]).all()
    assert np.isnan(Y[(sid < 0) | (sid > 2)]).all()
    ctx.close()


def test_ws_deep_nodes(torch_mod):
    """Nodes with 8 to 15 factors (16-row path stack) on the WS sweep."""
    from test_gpu_deep import deep_node_set
    ii, jj, lv = deep_node_set(8, 3000)
    S = hddm_lb.synthetic_surplus(lv, 96, 9)
    nact = ((ii != 0) | (jj != 0)).sum(axis=1)
    S[nact >= 8] *= 2.0 ** 14
    X = hddm_lb.synthetic_points(500, 16, 10, 0.02, 0.45)
    ctx = ws_ctx()
    ctx.upload_state(0, ii, jj, S)
    yo = oracle.forward(ii, jj, S, X)
    for prec in ALL_PRECISIONS:
        y = run_forward(ctx, X, precision=prec)
        assert ctx.device_status() == 0
        err = normwise_rel_err(y, yo)
        say("ws deep nodes precision", prec, "err", err)
        assert err <= TOL[prec]
    ctx.close()


def test_ws_matches_classic_sweep(torch_mod):
    """Same inputs through both formulations: FP64 agrees to rounding, and
    the point order of the output does not depend on the formulation."""
    ii, jj, lv, _ = hddm_lb.generate_grid(40, 2, 4, 0.3, 90, 8000)
    S = hddm_lb.synthetic_surplus(lv, 128, 91)
    X = hddm_lb.synthetic_points(2000, 40, 92)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    ctx.set_option(hddm_lb.OPT_FORMULATION, 1)
    y1 = run_forward(ctx, X, precision=hddm_lb.FP64)
    ctx.set_option(hddm_lb.OPT_FORMULATION, 2)
    y2 = run_forward(ctx, X, precision=hddm_lb.FP64)
    assert ctx.device_status() == 0
    assert normwise_rel_err(y2, y1) <= 1e-13
    ctx.close()


def test_ws_stats_count_nonzero_pairs(torch_mod):
    """Statistics mode of the WS sweep: the non-zero (point, node) pairs it
    reports equal the count from the oracle's basis."""
    torch = torch_mod
    ii, jj, lv, _ = hddm_lb.generate_grid(6, 2, 4, 0.5, 5, 0)
    nno = ii.shape[0]
    S = hddm_lb.synthetic_surplus(lv, 64, 6)
    X = hddm_lb.synthetic_points(257, 6, 7)
    ctx = ws_ctx()
    ctx.upload_state(0, ii, jj, S)
    ctx.set_option(hddm_lb.OPT_STATS, 1)
    run_forward(ctx, X, precision=hddm_lb.F32)
    st = ctx.last_stats(6)
    B = oracle.forward(ii, jj, np.eye(nno), X)
    say("ws stats", st, "nnz oracle", int((B != 0).sum()))
    assert st[5] == int((B != 0).sum())
    assert st[3] > 0 and st[4] > 0
    ctx.close()
