"""Multistate parity at the shapes of the batched workloads.

Multistate forward, one persistent launch: 12 states, D=20, adaptive grids
(max level 6) with node counts log-uniform in [2k, 200k], dofs per state from
{16, 32, 64}, 262,144 points with seeded state ids. High-dimensional
multistate: 4 states, D=59, regular grids sum(l-1) <= 3 (281,077 nodes),
dofs=128, 65,536 points (the 2^20-point workload's per-point work is the same;
the point count only scales time). The oracle evaluates a seeded subset of
points of every state; the whole batch is checked for finiteness and for
rows of invalid state ids staying untouched.
"""
import numpy as np
import pytest

import hddm_lb
import oracle
from gpu_util import TOL, normwise_rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def build_states(ctx, specs):
    states = []
    for s, (dim, start, maxl, p, seed, target, dofs) in enumerate(specs):
        ii, jj, lv, _ = hddm_lb.generate_grid(dim, start, maxl, p, seed, target)
        S = hddm_lb.synthetic_surplus(lv, dofs, seed + 100)
        ctx.upload_state(s, ii, jj, S)
        states.append((ii, jj, S))
        say("state", s, "nno", ii.shape[0], "dofs", dofs)
    return states


def check_subset(states, X, sid, Y, prec, per_state, seed):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for s, (ii, jj, S) in enumerate(states):
        rows = np.flatnonzero(sid == s)
        pick = rng.choice(rows, size=min(per_state, rows.size), replace=False)
        yo = oracle.forward(ii, jj, S, X[pick])
        err = normwise_rel_err(Y[pick, :S.shape[1]], yo)
        worst = max(worst, err)
        assert err <= TOL[prec], (s, err)
    return worst


@pytest.fixture(scope="module")
def d20_twelve_states(torch_mod):
    rng = np.random.default_rng(2024)
    nnos = np.exp(rng.uniform(np.log(2000), np.log(200000), size=12)).astype(int)
    dofs = rng.choice([16, 32, 64], size=12)
    ctx = hddm_lb.Context(0)
    specs = [(20, 1, 6, 0.3, 300 + s, int(nnos[s]), int(dofs[s])) for s in range(12)]
    states = build_states(ctx, specs)
    n = 262144
    X = hddm_lb.synthetic_points(n, 20, 77)
    w = rng.uniform(0.5, 1.5, size=12)
    sid = rng.choice(12, size=n, p=w / w.sum()).astype(np.int32)
    yield ctx, states, X, sid
    ctx.close()


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32, hddm_lb.F16X2])
def test_multistate_d20_twelve_states(torch_mod, d20_twelve_states, prec):
    torch = torch_mod
    ctx, states, X, sid = d20_twelve_states
    x = torch.from_numpy(X).cuda()
    s = torch.from_numpy(sid).cuda()
    y = torch.full((X.shape[0], 64), np.nan, dtype=torch.float64, device="cuda")
    ctx.forward(x, y, s, prec)
    torch.cuda.synchronize()
    Y = y.cpu().numpy()
    for st, (_, _, S) in enumerate(states):
        rows = sid == st
        assert np.isfinite(Y[rows, :S.shape[1]]).all()
        assert np.isnan(Y[rows, S.shape[1]:]).all()
    worst = check_subset(states, X, sid, Y, prec, 24, 5)
    say("multistate D=20 precision", prec, "worst normwise err", worst)


@pytest.fixture(scope="module")
def d59_four_states(torch_mod):
    ctx = hddm_lb.Context(0)
    states = build_states(ctx, [(59, 3, 4, 0.0, 500 + s, 0, 128) for s in range(4)])
    n = 65536
    X = hddm_lb.synthetic_points(n, 59, 91)
    sid = np.random.default_rng(3).integers(0, 4, size=n).astype(np.int32)
    yield ctx, states, X, sid
    ctx.close()


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32, hddm_lb.F16X2, hddm_lb.BF16X3])
def test_multistate_d59_dofs128(torch_mod, d59_four_states, prec):
    torch = torch_mod
    ctx, states, X, sid = d59_four_states
    x = torch.from_numpy(X).cuda()
    s = torch.from_numpy(sid).cuda()
    y = torch.full((X.shape[0], 128), np.nan, dtype=torch.float64, device="cuda")
    ctx.forward(x, y, s, prec)
    torch.cuda.synchronize()
    Y = y.cpu().numpy()
    assert np.isfinite(Y).all()
    worst = check_subset(states, X, sid, Y, prec, 12, 6)
    say("multistate D=59 precision", prec, "worst normwise err", worst)
