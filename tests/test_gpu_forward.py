"""GPU parity of the forward sweep against the CPU oracle (gold restatement).

Tolerances (normwise per dof column, max |dy| / max |y_gold|): FP64 <= 1e-12,
every mixed-precision MFMA path <= 1e-6, as the target states.
"""
import json
import os
import sys

import numpy as np
import pytest

import hddm_lb
import oracle
from conftest import GOLDEN
from gpu_util import ALL_PRECISIONS, TOL, normwise_rel_err, run_forward

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    c = hddm_lb.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)


def test_kat_on_device(kat):
    """The reference's known answer (LinearBasis/tests/InterpolateArray/test.cpp:
    118 values within EPSILON = 1e-3) on every path, and parity with the oracle.

    This single point is ill-conditioned in some dof columns (sum |B S| / |y|
    up to 17), so the per-column error of one point is the strictest check
    here. FP64 and the F32 MFMA path meet their tolerances per column. The
    f16 and bf16 split paths round each chunk's partial sum to FP32 inside
    the MFMA, which on this point costs up to ~1.6e-6 in a cancelling column:
    they are held to 1e-6 normwise over the state (max |dy| / max |y|) and to
    3e-6 per column."""
    st = kat["state"]
    c = hddm_lb.Context(0)
    c.upload_state(0, st["index_i"], st["index_j"], st["surplus"])
    X = np.array([kat["x"]])
    yo = oracle.forward(st["index_i"], st["index_j"], st["surplus"], X)
    for prec in ALL_PRECISIONS:
        y = run_forward(c, X, precision=prec)
        col, state = normwise_rel_err(y, yo), np.abs(y - yo).max() / np.abs(yo).max()
        say("kat precision", prec, "per-column err", col, "per-state err", state)
        np.testing.assert_allclose(y[0], kat["expected"], rtol=0, atol=kat["epsilon"])
        if prec in (hddm_lb.FP64, hddm_lb.F32):
            assert col <= TOL[prec], prec
        else:
            assert state <= TOL[prec] and col <= 3e-6, prec
    c.close()


@pytest.mark.parametrize("prec", ALL_PRECISIONS)
def test_regular_grid_d4_401_nodes(ctx, prec):
    """Regular D=4 grid sum(l-1) <= 4 (401 nodes), dofs=8, 4096 points."""
    ii, jj, lv, ix = hddm_lb.generate_grid(4, 4, 5, 1.0, 1, 0)
    S = hddm_lb.synthetic_surplus(lv, 8, 11)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(4096, 4, 12)
    say("regular d4 precision", prec)
    y = run_forward(ctx, X, precision=prec)
    yo = oracle.forward(ii, jj, S, X)
    err = normwise_rel_err(y, yo)
    say("  normwise rel err", err)
    assert err <= TOL[prec]


@pytest.mark.parametrize("prec", ALL_PRECISIONS)
def test_adaptive_grid_d16_60k_subset(ctx, prec):
    """D=16 adaptive grid of ~60k nodes, dofs=32, on a 1536-point subset."""
    ii, jj, lv, ix = hddm_lb.generate_grid(16, 2, 7, 0.3, 1, 60000)
    S = hddm_lb.synthetic_surplus(lv, 32, 2)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(1536, 16, 3)
    say("d16 60k subset precision", prec)
    y = run_forward(ctx, X, precision=prec)
    yo = oracle.forward(ii, jj, S, X)
    err = normwise_rel_err(y, yo)
    say("  normwise rel err", err)
    assert err <= TOL[prec]


@pytest.mark.parametrize("dim", [15, 16, 17, 24, 32])
def test_coordinate_bank_edges(ctx, dim):
    """Dimensions at and around the one/two coordinate-bank boundary (the
    walk reads x through a register index of twice the dimension), every
    precision, against the oracle."""
    ii, jj, lv, ix = hddm_lb.generate_grid(dim, 2, 5, 0.4, 40 + dim, 4000)
    S = hddm_lb.synthetic_surplus(lv, 32, 41 + dim)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(700, dim, 42 + dim)
    yo = oracle.forward(ii, jj, S, X)
    for prec in ALL_PRECISIONS:
        say("bank edge dim", dim, "nno", ii.shape[0], "precision", prec)
        y = run_forward(ctx, X, precision=prec)
        assert normwise_rel_err(y, yo) <= TOL[prec]


@pytest.mark.parametrize("n", [1, 2, 63, 65, 200])
def test_ragged_point_counts(ctx, n):
    ii, jj, lv, ix = hddm_lb.generate_grid(6, 2, 5, 0.5, 5, 3000)
    S = hddm_lb.synthetic_surplus(lv, 20, 6)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(n, 6, 7)
    yo = oracle.forward(ii, jj, S, X)
    for prec in (hddm_lb.FP64, hddm_lb.F16X2):
        say("ragged n", n, "precision", prec)
        y = run_forward(ctx, X, precision=prec)
        assert normwise_rel_err(y, yo) <= TOL[prec]


def test_points_on_support_edges(ctx):
    ii, jj, lv, ix = hddm_lb.generate_grid(5, 3, 5, 0.6, 8, 2000)
    S = hddm_lb.synthetic_surplus(lv, 6, 9)
    ctx.upload_state(0, ii, jj, S)
    rng = np.random.default_rng(1)
    grid_vals = np.array([0.0, 1.0, 0.5, 0.25, 0.75, 0.125, 0.875, 1.0 / 3])
    X = grid_vals[rng.integers(0, len(grid_vals), size=(300, 5))]
    yo = oracle.forward(ii, jj, S, X)
    for prec in ALL_PRECISIONS:
        say("edges precision", prec)
        y = run_forward(ctx, X, precision=prec)
        assert normwise_rel_err(y, yo) <= TOL[prec]


def test_fp64_matches_gold_basis_bitwise_single_node(ctx):
    """One node per grid (plus root): y = S_root + B_k S_k computed with gold's
    operation order must agree to the last bit in FP64."""
    ii = np.array([[0, 0, 0], [4, 8, 2]], np.int32)
    jj = np.array([[0, 0, 0], [1, 5, 0]], np.int32)
    S = np.array([[0.0], [1.0]])
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(256, 3, 4)
    y = run_forward(ctx, X, precision=hddm_lb.FP64)
    yo = oracle.forward(ii, jj, S, X)
    assert np.array_equal(y, yo)


def test_multistate_mixed_dofs(ctx):
    """Several states, per-state node counts and dofs (16, 32, 64, 40), points
    assigned by a seeded categorical draw, one launch."""
    c = hddm_lb.Context(0)
    specs = [(300, 16), (2500, 32), (800, 64), (5000, 40), (150, 16)]
    states = []
    for s, (target, dofs) in enumerate(specs):
        ii, jj, lv, ix = hddm_lb.generate_grid(8, 2, 6, 0.3, 100 + s, target)
        S = hddm_lb.synthetic_surplus(lv, dofs, 200 + s)
        c.upload_state(s, ii, jj, S)
        states.append((ii, jj, S))
    rng = np.random.default_rng(5)
    n = 3000
    sid = rng.choice(len(specs), size=n, p=[0.1, 0.3, 0.2, 0.3, 0.1]).astype(np.int32)
    X = hddm_lb.synthetic_points(n, 8, 6)
    for prec in ALL_PRECISIONS:
        say("multistate precision", prec)
        y = run_forward(c, X, sid, precision=prec, ldy=64)
        for s, (ii, jj, S) in enumerate(states):
            m = sid == s
            yo = oracle.forward(ii, jj, S, X[m])
            assert normwise_rel_err(y[m, :S.shape[1]], yo) <= TOL[prec], (prec, s)
            if S.shape[1] < 64:
                assert np.all(np.isnan(y[m, S.shape[1]:]))   # beyond dofs: untouched
    c.close()


def test_invalid_state_ids_leave_rows_untouched(ctx):
    ii, jj, lv, ix = hddm_lb.generate_grid(4, 3, 4, 0.5, 1, 0)
    S = hddm_lb.synthetic_surplus(lv, 4, 1)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(100, 4, 2)
    sid = np.zeros(100, np.int32)
    sid[::3] = 7
    sid[1::5] = -1
    y = run_forward(ctx, X, sid, precision=hddm_lb.FP64)
    bad = (sid != 0)
    assert np.all(np.isnan(y[bad]))
    yo = oracle.forward(ii, jj, S, X[~bad])
    assert normwise_rel_err(y[~bad], yo) <= 1e-12


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F16X2])
def test_node_segments_match_unsplit(ctx, prec):
    """Work items split over node segments (partials + fixed-order reduce) agree with the oracle."""
    c = hddm_lb.Context(0)
    ii, jj, lv, ix = hddm_lb.generate_grid(10, 2, 6, 0.3, 4, 20000)
    S = hddm_lb.synthetic_surplus(lv, 48, 5)
    c.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(700, 10, 6)
    yo = oracle.forward(ii, jj, S, X)
    y1 = run_forward(c, X, precision=prec)
    c.set_option(hddm_lb.OPT_SEGMENT_NODES, 1000)
    say("segments precision", prec)
    y2 = run_forward(c, X, precision=prec)
    assert normwise_rel_err(y1, yo) <= TOL[prec]
    assert normwise_rel_err(y2, yo) <= TOL[prec]
    c.close()


def test_results_independent_of_point_order(ctx):
    """Per-point results do not depend on which points share a wave."""
    ii, jj, lv, ix = hddm_lb.generate_grid(12, 2, 6, 0.3, 9, 8000)
    S = hddm_lb.synthetic_surplus(lv, 32, 1)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(2000, 12, 2)
    perm = np.random.default_rng(0).permutation(2000)
    for prec in (hddm_lb.FP64, hddm_lb.F16X2):
        a = run_forward(ctx, X, precision=prec)
        b = run_forward(ctx, X[perm], precision=prec)
        assert np.array_equal(a[perm], b)


def test_graph_plan_replays_match_eager(ctx):
    import torch
    ii, jj, lv, ix = hddm_lb.generate_grid(8, 2, 6, 0.3, 3, 5000)
    S = hddm_lb.synthetic_surplus(lv, 32, 3)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(5000, 8, 4)
    x = torch.from_numpy(X).cuda()
    eager = run_forward(ctx, X, precision=hddm_lb.F16X2)
    y = torch.full((5000, 32), np.nan, dtype=torch.float64, device="cuda")
    stream = torch.cuda.Stream()
    plan = ctx.plan(x, y, None, hddm_lb.F16X2, stream)
    for rep in range(4):
        say("graph replay", rep)
        y.fill_(np.nan)
        torch.cuda.synchronize()
        plan.launch(stream)
        stream.synchronize()
        assert np.array_equal(y.cpu().numpy(), eager)
        _ = run_forward(ctx, X[:100], precision=hddm_lb.FP64)  # eager work between replays
    del plan


def test_forward_host_path(ctx, kat):
    st = kat["state"]
    c = hddm_lb.Context(0)
    c.upload_state(0, st["index_i"], st["index_j"], st["surplus"])
    y = np.zeros((1, 118))
    c.forward_host(np.array([kat["x"]]), y, precision=hddm_lb.FP64)
    np.testing.assert_allclose(y[0], kat["expected"], rtol=0, atol=kat["epsilon"])
    c.close()


@pytest.mark.parametrize("walk,waves", [(1, 1), (1, 4), (2, 1), (2, 2), (2, 4)])
def test_walk_modes_and_group_sizes(ctx, walk, waves):
    """Both op-stream sources (scalar loads per wave, LDS tiles per workgroup)
    and every workgroup size give the oracle's results, with and without node
    segments (several LDS tiles per segment, ragged last group)."""
    ii, jj, lv, ix = hddm_lb.generate_grid(10, 2, 6, 0.4, 21, 9000)
    S = hddm_lb.synthetic_surplus(lv, 40, 22)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(1000, 10, 23)
    yo = oracle.forward(ii, jj, S, X)
    ctx.set_option(hddm_lb.OPT_WALK, walk)
    ctx.set_option(hddm_lb.OPT_WAVES_PER_WG, waves)
    try:
        for seg in (0, 2500):
            ctx.set_option(hddm_lb.OPT_SEGMENT_NODES, seg)
            for prec in (hddm_lb.FP64, hddm_lb.F16X2):
                say("walk", walk, "waves", waves, "segment", seg, "precision", prec)
                y = run_forward(ctx, X, precision=prec)
                assert normwise_rel_err(y, yo) <= TOL[prec]
    finally:
        ctx.set_option(hddm_lb.OPT_WALK, 0)
        ctx.set_option(hddm_lb.OPT_WAVES_PER_WG, 0)
        ctx.set_option(hddm_lb.OPT_SEGMENT_NODES, 0)


def test_stats_counters(ctx):
    """Statistics mode leaves the results bit for bit unchanged and reports
    consistent counts: the F32 path issues at most 2 MFMAs (two 32-point
    halves) per K-step of each 32-node chunk, the FP64 path none."""
    ii, jj, lv, ix = hddm_lb.generate_grid(12, 2, 6, 0.3, 31, 6000)
    S = hddm_lb.synthetic_surplus(lv, 32, 32)
    ctx.upload_state(0, ii, jj, S)
    X = hddm_lb.synthetic_points(3000, 12, 33)
    try:
        for prec in (hddm_lb.F32, hddm_lb.FP64):
            say("stats precision", prec)
            a = run_forward(ctx, X, precision=prec)
            ctx.set_option(hddm_lb.OPT_STATS, 1)
            b = run_forward(ctx, X, precision=prec)
            ctx.set_option(hddm_lb.OPT_STATS, 0)
            assert np.array_equal(a, b)
            ops, prunes, chunks, items, mfma = ctx.last_stats(5)
            assert ctx.last_stats() == (ops, prunes, chunks, items)
            assert items >= (3000 + 63) // 64 and ops > 0 and chunks > 0 and prunes < ops
            if prec == hddm_lb.F32:
                assert 0 < mfma <= 2 * 16 * chunks
            else:
                assert mfma == 0
        with pytest.raises(hddm_lb.HddmLbError):
            ctx.last_stats(7)
    finally:
        ctx.set_option(hddm_lb.OPT_STATS, 0)


def test_padded_rows(ctx):
    """Row strides larger than dim and dofs: padding columns of x hold NaN
    (any read of them would poison the result), padding columns of y are left
    untouched, and results equal those of contiguous rows bit for bit."""
    import torch
    ii, jj, lv, ix = hddm_lb.generate_grid(12, 2, 6, 0.3, 41, 3000)
    S = hddm_lb.synthetic_surplus(lv, 24, 42)
    ctx.upload_state(0, ii, jj, S)
    n, dim, dofs = 777, 12, 24
    X = hddm_lb.synthetic_points(n, dim, 43)
    xp = np.full((n, dim + 5), np.nan)
    xp[:, :dim] = X
    x = torch.from_numpy(xp).cuda()
    for prec in (hddm_lb.FP64, hddm_lb.F32):
        say("padded rows precision", prec)
        ref = run_forward(ctx, X, precision=prec)
        y = torch.full((n, dofs + 3), np.nan, dtype=torch.float64, device="cuda")
        ctx.forward(x, y, None, prec)
        torch.cuda.synchronize()
        yh = y.cpu().numpy()
        assert np.array_equal(yh[:, :dofs], ref)
        assert np.isnan(yh[:, dofs:]).all()


def test_forward_host_small_batches_replay_graphs(ctx):
    """Host-pointer calls of up to 64 points replay a captured forward: sizes
    on both sides of that bound, repeated calls, state ids with invalid rows,
    a padded y, and a state re-uploaded between calls (the captured forward
    is rebuilt)."""
    ii, jj, lv, _ = hddm_lb.generate_grid(9, 2, 5, 0.4, 21, 3000)
    S = hddm_lb.synthetic_surplus(lv, 24, 22)
    S2 = hddm_lb.synthetic_surplus(lv, 24, 23)
    c = hddm_lb.Context(0)
    c.upload_state(0, ii, jj, S)
    for n in (1, 3, 64, 65, 1, 64):
        X = hddm_lb.synthetic_points(n, 9, 100 + n)
        sid = np.zeros(n, np.int32)
        sid[1::3] = 5   # not loaded: row left as passed
        y = np.full((n, 30), np.nan)
        c.forward_host(X, y, sid, hddm_lb.FP64)
        yo = oracle.forward(ii, jj, S, X)
        ok = sid == 0
        assert np.isnan(y[~ok]).all() and np.isnan(y[:, 24:]).all()
        assert normwise_rel_err(y[ok, :24], yo[ok]) <= 1e-12, n
    c.upload_state(0, ii, jj, S2)
    X = hddm_lb.synthetic_points(3, 9, 103)
    y = np.full((3, 24), np.nan)
    c.forward_host(X, y, None, hddm_lb.F32)
    assert normwise_rel_err(y, oracle.forward(ii, jj, S2, X)) <= 1e-6
    c.close()


def test_autotuner_picks_and_caches(ctx):
    """With the autotuner on, the first forward of a key times the sweep
    formulations and keeps one; later forwards of the key come from the
    cache; every result matches the oracle."""
    ii, jj, lv, _ = hddm_lb.generate_grid(40, 2, 4, 0.3, 31, 6000)
    S = hddm_lb.synthetic_surplus(lv, 96, 32)
    X = hddm_lb.synthetic_points(3000, 40, 33)
    c = hddm_lb.Context(0)
    c.set_option(hddm_lb.OPT_AUTOTUNE, 1)
    c.upload_state(0, ii, jj, S)
    pick = np.random.default_rng(1).choice(3000, size=64, replace=False)
    yo = oracle.forward(ii, jj, S, X[pick])
    for rep in range(3):
        for prec in (hddm_lb.F32, hddm_lb.FP64):
            y = run_forward(c, X, precision=prec)
            assert normwise_rel_err(y[pick], yo) <= TOL[prec]
    cfg = c.last_config()
    say("autotuner config", cfg)
    assert cfg["tuned"] == 2 and cfg["cache_hits"] == 4
    assert cfg["formulation"] in (1, 2)
    assert c.device_status() == 0
    c.close()


def test_forward_host_threads_coalesce(ctx):
    """Eight host threads each make 60 one-point calls on one context (two
    states of different dofs, some invalid ids, a padded y): every result
    matches the oracle and invalid rows stay as passed; with coalescing on,
    the calls are served by fewer batched forwards than calls, and with it off
    they still give the same results."""
    import threading
    ii, jj, lv, _ = hddm_lb.generate_grid(7, 2, 5, 0.4, 81, 2000)
    S0 = hddm_lb.synthetic_surplus(lv, 12, 82)
    S1 = hddm_lb.synthetic_surplus(lv, 20, 83)
    c = hddm_lb.Context(0)
    c.upload_state(0, ii, jj, S0)
    c.upload_state(1, ii, jj, S1)
    nthreads, calls = 8, 60
    X = hddm_lb.synthetic_points(nthreads * calls, 7, 84)
    sid = np.random.default_rng(85).integers(-1, 2, size=X.shape[0]).astype(np.int32)
    ref = np.full((X.shape[0], 24), np.nan)
    for s, S in enumerate((S0, S1)):
        rows = sid == s
        ref[rows, :S.shape[1]] = oracle.forward(ii, jj, S, X[rows])
    for coalesce in (1, 0):
        c.set_option(hddm_lb.OPT_COALESCE, coalesce)
        out = np.full((X.shape[0], 24), np.nan)
        errors = []

        def worker(t):
            try:
                for k in range(calls):
                    p = t * calls + k
                    y = np.full((1, 24), np.nan)
                    c.forward_host(X[p:p + 1], y, sid[p:p + 1], hddm_lb.FP64)
                    out[p] = y[0]
            except Exception as e:  # reported below
                errors.append(e)
        before = c.host_call_stats()
        th = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        after = c.host_call_stats()
        assert not errors, errors
        ok = ~np.isnan(ref)
        assert np.array_equal(np.isnan(out), np.isnan(ref))
        assert np.abs(out[ok] - ref[ok]).max() <= 1e-12 * np.abs(ref[ok]).max()
        reqs, batches = after[0] - before[0], after[1] - before[1]
        say("coalesce", coalesce, "requests", reqs, "batches", batches)
        if coalesce:
            assert reqs == nthreads * calls and 0 < batches <= reqs
        else:
            assert reqs == 0
    c.close()
