"""GPU parity for nodes with 8 to 15 active factors (the 16-row path stack).

The sweep and adjoint kernels keep the prefix products of the current trie
path in a per-wave LDS stack of 8 rows when every node has at most 7 active
factors and 16 rows otherwise. The seeded grids of the other tests stay below
8 factors, so this node set is built directly: the regular grid
sum(l-1) <= 2 in D=16 plus random nodes with 8 to 15 active dims at levels 2
and 3 (a node set needs no hierarchical closure: the engine interpolates
whatever nodes it is given). Points lie in [0.02, 0.45]^16, where the
level-2 factors with j = 0 and the level-3 factors with j = 1 are non-zero, so
deep nodes survive the walk; nodes with a j = 2 or j = 3 factor are zero
there and exercise the pruning.
"""
import numpy as np
import pytest

import hddm_lb
import oracle
from gpu_util import ALL_PRECISIONS, TOL, normwise_rel_err, run_forward

pytestmark = pytest.mark.gpu

DIM = 16


def say(*a):
    print(*a, flush=True)


def deep_node_set(seed, ndeep):
    ii, jj, lv, _ = hddm_lb.generate_grid(DIM, 2, 3, 0.0, seed, 0)
    rng = np.random.default_rng(seed)
    di = np.zeros((ndeep, DIM), np.int32)
    dj = np.zeros((ndeep, DIM), np.int32)
    dl = np.ones((ndeep, DIM), np.int32)
    for k in range(ndeep):
        nact = int(rng.integers(8, 16))
        dims = rng.choice(DIM, size=nact, replace=False)
        for d in dims:
            if rng.random() < 0.6:
                di[k, d], dj[k, d], dl[k, d] = 2, (0 if rng.random() < 0.9 else 2), 2
            else:
                di[k, d], dj[k, d], dl[k, d] = 4, (1 if rng.random() < 0.9 else 3), 3
    ii = np.concatenate([ii, di])
    jj = np.concatenate([jj, dj])
    lv = np.concatenate([lv, dl])
    _, keep = np.unique(np.concatenate([ii, jj], axis=1), axis=0, return_index=True)
    keep = np.sort(keep)
    return ii[keep], jj[keep], lv[keep]


@pytest.fixture(scope="module")
def deep_case():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    ii, jj, lv = deep_node_set(8, 3000)
    S = hddm_lb.synthetic_surplus(lv, 40, 9)
    nact = ((ii != 0) | (jj != 0)).sum(axis=1)
    # 2^-sum(l) makes deep surpluses tiny; scale them so they carry weight
    S[nact >= 8] *= 2.0 ** 14
    n = 700
    X = hddm_lb.synthetic_points(n, DIM, 10, 0.02, 0.45, kink_level=3, kink_gap=1e-3)
    assert nact.max() >= 12 and (nact >= 8).sum() > 1000
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    assert ctx.state_info(0)["max_factors"] >= 8
    yield ctx, ii, jj, S, X
    ctx.close()


@pytest.mark.parametrize("prec", ALL_PRECISIONS)
def test_deep_nodes_forward(deep_case, prec):
    ctx, ii, jj, S, X = deep_case
    y = run_forward(ctx, X, precision=prec)
    yo = oracle.forward(ii, jj, S, X)
    # the deep nodes must carry weight at these points, or the test shows nothing
    deep = ((ii != 0) | (jj != 0)).sum(axis=1) >= 8
    yd = oracle.forward(ii[deep], jj[deep], S[deep], X[:50])
    assert np.abs(yd).max() > 0.1 * np.abs(yo[:50]).max()
    err = normwise_rel_err(y, yo)
    say("deep nodes precision", prec, "err", err)
    assert err <= TOL[prec]


@pytest.mark.parametrize("form", [0, 2])
def test_deep_nodes_adjoint(deep_case, form):
    """Automatic choice (the per-tile kernels at D=16) and the
    wave-specialized kernel (16 factor slots per node in its dx wave)."""
    import torch
    ctx, ii, jj, S, X = deep_case
    ctx.set_option(hddm_lb.OPT_FORMULATION, form)
    G = np.random.default_rng(11).standard_normal((X.shape[0], S.shape[1]))
    x = torch.from_numpy(X).cuda()
    g = torch.from_numpy(G).cuda()
    dx = torch.full(X.shape, np.nan, dtype=torch.float64, device="cuda")
    _, total = ctx.surplus_offset(-1)
    ds = torch.full((total,), np.nan, dtype=torch.float64, device="cuda")
    ctx.adjoint(x, g, dx, ds, None, hddm_lb.FP64)
    torch.cuda.synchronize()
    ctx.set_option(hddm_lb.OPT_FORMULATION, 0)
    assert ctx.device_status() == 0
    dxo, dso = oracle.adjoint(ii, jj, S, X, G)
    ex = normwise_rel_err(dx.cpu().numpy(), dxo)
    es = normwise_rel_err(ds.cpu().numpy().reshape(dso.shape), dso)
    say("deep nodes adjoint, formulation", form, "dx err", ex, "ds err", es)
    assert ex <= 1e-10 and es <= 1e-12
