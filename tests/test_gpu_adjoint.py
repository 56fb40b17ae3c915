"""GPU parity of the adjoint (dx, dsurplus) through the C ABI.

Two kernels compute it: the wave-specialized one (a walker wave and MFMA
consumer waves, FP64 or F32; HDDM_LB_OPT_FORMULATION = 2) and the per-tile
pair of FP64 kernels (= 1, FP64 arithmetic for either precision). F32 is held
to 1e-6 normwise, FP64 to the bounds below. Every call must leave the device
status word clear.

Checks, in FP64:
  * against the oracle's analytic adjoint (gold basis, the product of the
    other factors formed directly): dx and dS to 1e-10 normwise;
  * dx against central finite differences of the oracle forward (h = 1e-4,
    points at least 1e-3 from every kink, where the interpolant is linear in
    each coordinate, so the difference quotient is exact up to rounding);
  * dS against B^T g written out from the oracle's basis on a tiny grid;
  * the torch.autograd.Function harness on the same inputs.
On the per-tile kernels and on the wave-specialized kernel without owned
units the dsurplus sums use FP64 atomics; their order varies, hence 1e-12
and not bitwise. With owned units (one workgroup adds to each row of one of
K dS copies; HDDM_LB_OPT_ADJOINT_OWNED) dsurplus is bitwise repeatable.
"""
import numpy as np
import pytest

import hddm_lb
import oracle
from gpu_util import normwise_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def say(*a):
    print(*a, flush=True)


# (dx, dS) normwise bounds per precision
TOL_ADJ = {hddm_lb.FP64: (1e-10, 1e-12), hddm_lb.F32: (1e-6, 1e-6)}
# (precision, formulation): FP64 on both kernels, F32 on the wave-specialized
# one and on the automatic choice
ADJ_PATHS = [(hddm_lb.FP64, 2), (hddm_lb.FP64, 1), (hddm_lb.F32, 2), (hddm_lb.F32, 0)]


def run_adjoint(torch, ctx, X, G, sid=None, prec=hddm_lb.FP64):
    x = torch.from_numpy(X).cuda()
    g = torch.from_numpy(G).cuda()
    dx = torch.full(X.shape, np.nan, dtype=torch.float64, device="cuda")
    _, total = ctx.surplus_offset(-1)
    ds = torch.full((total,), np.nan, dtype=torch.float64, device="cuda")
    s = None if sid is None else torch.from_numpy(np.asarray(sid, np.int32)).cuda()
    ctx.adjoint(x, g, dx, ds, s, prec)
    torch.cuda.synchronize()
    assert ctx.device_status() == 0
    return dx.cpu().numpy(), ds.cpu().numpy()


def adj_ctx(form=0):
    ctx = hddm_lb.Context(0)
    if form:
        ctx.set_option(hddm_lb.OPT_FORMULATION, form)
    ctx.device_status(clear=True)
    return ctx


def grid_case(dim, start, maxl, p, seed, target, dofs, n, kink_level):
    ii, jj, lv, _ = hddm_lb.generate_grid(dim, start, maxl, p, seed, target)
    S = hddm_lb.synthetic_surplus(lv, dofs, seed + 1)
    X = hddm_lb.synthetic_points(n, dim, seed + 2, 0.01, 0.99, kink_level=kink_level, kink_gap=1e-3)
    G = np.random.default_rng(seed + 3).standard_normal((n, dofs))
    return ii, jj, S, X, G


@pytest.mark.parametrize("prec,form", ADJ_PATHS)
@pytest.mark.parametrize("dim,dofs,n", [(6, 20, 300), (16, 32, 200), (17, 16, 150), (20, 40, 130), (32, 32, 120),
                                         (33, 24, 110), (59, 64, 100), (64, 40, 100), (12, 160, 90)])
def test_adjoint_matches_oracle(torch_mod, dim, dofs, n, prec, form):
    """One, two and four coordinate banks (D <= 16, <= 32, <= 64); one to
    five dof tiles (two launches of the wave-specialized kernel)."""
    ii, jj, S, X, G = grid_case(dim, 2, 5, 0.4, 3 + dim, 2500, dofs, n, 5)
    ctx = adj_ctx(form)
    ctx.upload_state(0, ii, jj, S)
    say("adjoint dim", dim, "nno", ii.shape[0], "dofs", dofs, "n", n, "precision", prec, "formulation", form)
    dx, ds = run_adjoint(torch_mod, ctx, X, G, prec=prec)
    dxo, dso = oracle.adjoint(ii, jj, S, X, G)
    ds = ds.reshape(ii.shape[0], dofs)
    say("  dx err", normwise_rel_err(dx, dxo), "ds err", normwise_rel_err(ds, dso))
    assert np.isfinite(dx).all() and np.isfinite(ds).all()
    assert normwise_rel_err(dx, dxo) <= TOL_ADJ[prec][0]
    assert normwise_rel_err(ds, dso) <= TOL_ADJ[prec][1]
    ctx.close()


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32])
@pytest.mark.parametrize("start,maxl,maxf", [(5, 6, 5), (6, 7, 6)])
def test_adjoint_five_and_six_factor_nodes(torch_mod, start, maxl, maxf, prec):
    """Regular D=6 grids whose nodes have up to 5 and 6 factors: the
    wave-specialized kernel's dx waves read 5 and 6 of the 8 factor slots.
    Against the oracle, 64 dofs (two tiles)."""
    ii, jj, lv, _ = hddm_lb.generate_grid(6, start, maxl, 0.0, 1, 0)
    S = hddm_lb.synthetic_surplus(lv, 64, 2)
    X = hddm_lb.synthetic_points(200, 6, 3, 0.01, 0.99, kink_level=maxl, kink_gap=1e-3)
    G = np.random.default_rng(4).standard_normal((200, 64))
    ctx = adj_ctx(2)
    ctx.upload_state(0, ii, jj, S)
    assert ctx.state_info(0)["max_factors"] == maxf
    say("adjoint D=6 nno", ii.shape[0], "max factors", maxf, "precision", prec)
    dx, ds = run_adjoint(torch_mod, ctx, X, G, prec=prec)
    assert ctx.last_adjoint_config()["kernels"] == 2
    dxo, dso = oracle.adjoint(ii, jj, S, X, G, nthreads=8)
    ds = ds.reshape(ii.shape[0], 64)
    say("  dx err", normwise_rel_err(dx, dxo), "ds err", normwise_rel_err(ds, dso))
    assert normwise_rel_err(dx, dxo) <= TOL_ADJ[prec][0]
    assert normwise_rel_err(ds, dso) <= TOL_ADJ[prec][1]
    ctx.close()


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32])
def test_adjoint_index_not_power_of_two(torch_mod, prec):
    """index_i = 3 in dim 0 of some nodes (the engine takes any index whose
    FP64 low word is zero; sparse grids have powers of two only). The
    wave-specialized kernel's dx waves use the factor table's centres -j / i,
    which are exact only for powers of two, so the automatic choice runs the
    per-tile kernels (against the oracle), and forcing the wave-specialized
    kernel is refused."""
    ii, jj, lv, _ = hddm_lb.generate_grid(6, 3, 4, 0.0, 5, 0)
    ii = ii.copy()
    sel = ii[:, 0] == 4          # level-3 nodes in dim 0: i = 4, j = 1 or 3
    assert sel.any()
    ii[sel, 0] = 3
    S = hddm_lb.synthetic_surplus(lv, 64, 6)
    X = hddm_lb.synthetic_points(150, 6, 7, 0.01, 0.99)
    G = np.random.default_rng(8).standard_normal((150, 64))
    ctx = adj_ctx(0)
    ctx.upload_state(0, ii, jj, S)
    say("adjoint D=6 nno", ii.shape[0], "nodes with i = 3:", int(sel.sum()), "precision", prec)
    dx, ds = run_adjoint(torch_mod, ctx, X, G, prec=prec)
    assert ctx.last_adjoint_config()["kernels"] == 1
    dxo, dso = oracle.adjoint(ii, jj, S, X, G, nthreads=8)
    ds = ds.reshape(ii.shape[0], 64)
    say("  dx err", normwise_rel_err(dx, dxo), "ds err", normwise_rel_err(ds, dso))
    assert normwise_rel_err(dx, dxo) <= TOL_ADJ[prec][0]
    assert normwise_rel_err(ds, dso) <= TOL_ADJ[prec][1]
    ctx.set_option(hddm_lb.OPT_FORMULATION, 2)
    with pytest.raises(hddm_lb.HddmLbError, match="powers of two"):
        run_adjoint(torch_mod, ctx, X, G, prec=prec)
    ctx.close()


def test_adjoint_kernels_agree(torch_mod):
    """The wave-specialized FP64 adjoint and the per-tile FP64 kernels give
    the same dx and dS to rounding (four dof tiles, node segments on)."""
    ii, jj, S, X, G = grid_case(24, 2, 5, 0.4, 66, 12000, 128, 700, 5)
    out = []
    for form in (1, 2):
        ctx = adj_ctx(form)
        ctx.set_option(hddm_lb.OPT_SEGMENT_NODES, 3000)
        ctx.upload_state(0, ii, jj, S)
        out.append(run_adjoint(torch_mod, ctx, X, G))
        ctx.close()
    assert normwise_rel_err(out[1][0], out[0][0]) <= 1e-13
    assert normwise_rel_err(out[1][1].reshape(-1, 128), out[0][1].reshape(-1, 128)) <= 1e-13


@pytest.mark.parametrize("which", ["dx", "ds"])
def test_adjoint_single_output_matches_both(torch_mod, which):
    """Asking for one output runs only its kernel; it must equal the same
    output of the call that asks for both."""
    ii, jj, S, X, G = grid_case(20, 2, 5, 0.4, 77, 2500, 32, 200, 5)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    dx, ds = run_adjoint(torch_mod, ctx, X, G)
    x = torch_mod.from_numpy(X).cuda()
    g = torch_mod.from_numpy(G).cuda()
    _, total = ctx.surplus_offset(-1)
    if which == "dx":
        d1 = torch_mod.full(X.shape, np.nan, dtype=torch_mod.float64, device="cuda")
        ctx.adjoint(x, g, d1, None, None, hddm_lb.FP64)
        torch_mod.cuda.synchronize()
        assert np.array_equal(d1.cpu().numpy(), dx)
    else:
        d1 = torch_mod.full((total,), np.nan, dtype=torch_mod.float64, device="cuda")
        ctx.adjoint(x, g, None, d1, None, hddm_lb.FP64)
        torch_mod.cuda.synchronize()
        # dS is summed with atomics: the order (last bits) may differ
        assert normwise_rel_err(d1.cpu().numpy().reshape(-1, 32), ds.reshape(-1, 32)) <= 1e-14
    ctx.close()


def test_adjoint_dx_matches_central_differences(torch_mod):
    ii, jj, S, X, G = grid_case(5, 2, 5, 0.5, 41, 1500, 8, 40, 5)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    dx, _ = run_adjoint(torch_mod, ctx, X, G)
    h = 1e-4
    fd = np.zeros_like(X)
    for d in range(X.shape[1]):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, d] += h
        Xm[:, d] -= h
        yp = oracle.forward(ii, jj, S, Xp)
        ym = oracle.forward(ii, jj, S, Xm)
        fd[:, d] = ((yp - ym) * G).sum(axis=1) / (2 * h)
    say("fd err", normwise_rel_err(dx, fd))
    assert normwise_rel_err(dx, fd) <= 1e-8
    ctx.close()


def test_adjoint_ds_is_basis_transpose_times_g(torch_mod):
    # tiny grid: B from the oracle forward with unit surpluses, dS = B^T g
    ii, jj, S, X, G = grid_case(3, 2, 4, 0.5, 5, 0, 3, 70, 4)
    nno = ii.shape[0]
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    _, ds = run_adjoint(torch_mod, ctx, X, G)
    B = oracle.forward(ii, jj, np.eye(nno), X)  # [n, nno]
    np.testing.assert_allclose(ds.reshape(nno, 3), B.T @ G, rtol=0, atol=1e-13 * np.abs(B.T @ G).max())
    ctx.close()


def test_adjoint_multistate_offsets(torch_mod):
    cases = [grid_case(6, 2, 4, 0.5, 60 + s, 600 + 300 * s, d, 100, 4) for s, d in enumerate((8, 20, 33))]
    ctx = hddm_lb.Context(0)
    for s, (ii, jj, S, _, _) in enumerate(cases):
        ctx.upload_state(s, ii, jj, S)
    X = np.concatenate([c[3] for c in cases])
    sid = np.repeat(np.arange(3), 100).astype(np.int32)
    rng = np.random.default_rng(7)
    perm = rng.permutation(X.shape[0])
    X, sid = X[perm], sid[perm]
    G = rng.standard_normal((X.shape[0], 33))
    dx, ds = run_adjoint(torch_mod, ctx, X, G, sid)
    for s, (ii, jj, S, _, _) in enumerate(cases):
        sel = sid == s
        dofs = S.shape[1]
        dxo, dso = oracle.adjoint(ii, jj, S, X[sel], G[sel, :dofs])
        off, _ = ctx.surplus_offset(s)
        dss = ds[off:off + ii.shape[0] * dofs].reshape(ii.shape[0], dofs)
        say("state", s, "dx err", normwise_rel_err(dx[sel], dxo), "ds err", normwise_rel_err(dss, dso))
        assert normwise_rel_err(dx[sel], dxo) <= 1e-10
        assert normwise_rel_err(dss, dso) <= 1e-12
    ctx.close()


def owned_pair(torch, cases, X, G, sid, prec, seg_nodes, one_copy=False):
    """The wave-specialized adjoint twice on the same states: dS through
    owned units, then through atomics. Returns ((dx, ds) owned, (dx, ds) a
    second owned call, (dx, ds) atomics, the owned call's configuration).
    one_copy: the owned call must add to dsurplus itself (K = 1)."""
    out = []
    for owned in (1, 0):
        ctx = adj_ctx(2)
        ctx.set_option(hddm_lb.OPT_SEGMENT_NODES, seg_nodes)
        ctx.set_option(hddm_lb.OPT_ADJOINT_OWNED, owned)
        for s, (ii, jj, S) in enumerate(cases):
            ctx.upload_state(s, ii, jj, S)
        out.append(run_adjoint(torch, ctx, X, G, sid, prec))
        cfg = ctx.last_adjoint_config()
        say("  owned option", owned, cfg)
        assert cfg["kernels"] == 2
        if owned:
            if one_copy:
                assert cfg["owned_copies"] == 1 and cfg["ds_copies"] == 1
            else:
                assert cfg["owned_copies"] >= 2
            owned_cfg = cfg
            out.append(run_adjoint(torch, ctx, X, G, sid, prec))
        else:
            assert cfg["owned_copies"] == 0
        ctx.close()
    return out[0], out[1], out[2], owned_cfg


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32])
def test_adjoint_owned_units(torch_mod, prec):
    """dS through owned units (segments of ~100 nodes, so every workgroup
    owns several (state, segment, copy) units): dx equal to the atomic
    path's to the last bit (same kernel, same segments), dS equal to it to
    rounding, both against the oracle, and dS bitwise repeatable."""
    ii, jj, S, X, G = grid_case(40, 2, 5, 0.4, 88, 6000, 96, 2000, 5)
    say("owned units D=40 nno", ii.shape[0], "dofs 96, 2000 points, precision", prec)
    own, own2, atom, cfg = owned_pair(torch_mod, [(ii, jj, S)], X, G, None, prec, 100)
    assert np.isfinite(own[0]).all() and np.isfinite(own[1]).all()
    assert np.array_equal(own[0], atom[0])
    assert np.array_equal(own[1], own2[1]) and np.array_equal(own[0], own2[0])
    ds_o, ds_a = own[1].reshape(-1, 96), atom[1].reshape(-1, 96)
    assert normwise_rel_err(ds_o, ds_a) <= 1e-14
    dxo, dso = oracle.adjoint(ii, jj, S, X, G, nthreads=8)
    say("  dx err", normwise_rel_err(own[0], dxo), "ds err", normwise_rel_err(ds_o, dso))
    assert normwise_rel_err(own[0], dxo) <= TOL_ADJ[prec][0]
    assert normwise_rel_err(ds_o, dso) <= TOL_ADJ[prec][1]


def test_adjoint_owned_units_single_copy(torch_mod):
    """Owned units with one dS copy: ten D=48 states split at every
    top-level run (about 180 segments each) give every workgroup several
    (state, segment) units, so each unit adds to dsurplus itself and nothing
    is summed after. dS against the atomic path and the oracle, bitwise
    repeatable."""
    cases, Xs = [], []
    for s in range(10):
        ii, jj, S, X, _ = grid_case(48, 2, 5, 0.5, 300 + s, 4000, 64, 240, 5)
        cases.append((ii, jj, S))
        Xs.append(X)
    X = np.concatenate(Xs)
    sid = np.repeat(np.arange(10), 240).astype(np.int32)
    rng = np.random.default_rng(12)
    perm = rng.permutation(X.shape[0])
    X, sid = X[perm], sid[perm]
    G = rng.standard_normal((X.shape[0], 64))
    own, own2, atom, cfg = owned_pair(torch_mod, cases, X, G, sid, hddm_lb.FP64, 1, one_copy=True)
    assert np.array_equal(own[1], own2[1])
    assert np.array_equal(own[0], atom[0])
    assert normwise_rel_err(own[1].reshape(-1, 64), atom[1].reshape(-1, 64)) <= 1e-14
    for s in (0, 7):
        ii, jj, S = cases[s]
        sel = sid == s
        dxo, dso = oracle.adjoint(ii, jj, S, X[sel], G[sel])
        off = sum(c[0].shape[0] * 64 for c in cases[:s])
        dss = own[1][off:off + ii.shape[0] * 64].reshape(ii.shape[0], 64)
        say("  state", s, "dx err", normwise_rel_err(own[0][sel], dxo), "ds err", normwise_rel_err(dss, dso))
        assert normwise_rel_err(own[0][sel], dxo) <= 1e-10
        assert normwise_rel_err(dss, dso) <= 1e-12


def test_adjoint_owned_units_multistate(torch_mod):
    """Owned units over three states of different node counts and dofs 24,
    64 and 130 (five dof tiles: two launches, the second with one state
    only), points in random state order: every state's dS rows against the
    oracle and the atomic path; dS bitwise repeatable."""
    dofs = (24, 64, 130)
    cases, Xs = [], []
    for s, d in enumerate(dofs):
        ii, jj, S, X, _ = grid_case(36, 2, 5, 0.4, 120 + s, 1500 + 1500 * s, d, 400, 5)
        cases.append((ii, jj, S))
        Xs.append(X)
    X = np.concatenate(Xs)
    sid = np.repeat(np.arange(3), 400).astype(np.int32)
    rng = np.random.default_rng(9)
    perm = rng.permutation(X.shape[0])
    X, sid = X[perm], sid[perm]
    G = rng.standard_normal((X.shape[0], 130))
    own, own2, atom, cfg = owned_pair(torch_mod, cases, X, G, sid, hddm_lb.FP64, 60)
    assert np.array_equal(own[1], own2[1])
    assert np.array_equal(own[0], atom[0])
    for s, (ii, jj, S) in enumerate(cases):
        sel = sid == s
        d = S.shape[1]
        dxo, dso = oracle.adjoint(ii, jj, S, X[sel], G[sel, :d])
        off = sum(c[0].shape[0] * c[2].shape[1] for c in cases[:s])
        dss = own[1][off:off + ii.shape[0] * d].reshape(ii.shape[0], d)
        dsa = atom[1][off:off + ii.shape[0] * d].reshape(ii.shape[0], d)
        say("  state", s, "dx err", normwise_rel_err(own[0][sel], dxo), "ds err", normwise_rel_err(dss, dso))
        assert normwise_rel_err(dss, dsa) <= 1e-14
        assert normwise_rel_err(own[0][sel], dxo) <= 1e-10
        assert normwise_rel_err(dss, dso) <= 1e-12


def test_autograd_harness(torch_mod):
    torch = torch_mod
    import hddm_lb_autograd
    ii, jj, S, X, G = grid_case(8, 2, 4, 0.5, 77, 1200, 16, 150, 4)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    x = torch.from_numpy(X).cuda().requires_grad_(True)
    s = torch.from_numpy(S).cuda().requires_grad_(True)
    y = hddm_lb_autograd.interpolate(ctx, x, s)
    (y * torch.from_numpy(G).cuda()).sum().backward()
    dxo, dso = oracle.adjoint(ii, jj, S, X, G)
    assert normwise_rel_err(y.detach().cpu().numpy(), oracle.forward(ii, jj, S, X)) <= 1e-12
    assert normwise_rel_err(x.grad.cpu().numpy(), dxo) <= 1e-10
    assert normwise_rel_err(s.grad.cpu().numpy(), dso) <= 1e-12
    ctx.close()


def test_adjoint_d20_100k_nodes(torch_mod):
    """Differentiated workload shape: D=20 adaptive grid of ~100k nodes (max
    level 6), dofs=32, 65,536 points in [0.01, 0.99]^20 kept 1e-3 off every
    kink, gbar ~ N(0, 1). dx is checked against the oracle on a seeded subset
    of points and against central differences on a few of them; dS against
    the oracle's B^T g on a 2,048-point slice."""
    torch = torch_mod
    ii, jj, lv, _ = hddm_lb.generate_grid(20, 1, 6, 0.3, 404, 100000)
    S = hddm_lb.synthetic_surplus(lv, 32, 405)
    n = 65536
    X = hddm_lb.synthetic_points(n, 20, 406, 0.01, 0.99, kink_level=6, kink_gap=1e-3)
    G = np.random.default_rng(407).standard_normal((n, 32))
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    say("adjoint D=20 nno", ii.shape[0])
    dx, ds = run_adjoint(torch, ctx, X, G)
    assert np.isfinite(dx).all() and np.isfinite(ds).all()
    sub = np.random.default_rng(408).choice(n, size=96, replace=False)
    dxo, _ = oracle.adjoint(ii, jj, S, X[sub], G[sub])
    say("  dx err (subset)", normwise_rel_err(dx[sub], dxo))
    assert normwise_rel_err(dx[sub], dxo) <= 1e-10
    # dS over a 2,048-point slice of the batch (the oracle's cost grows with the points)
    _, ds2 = run_adjoint(torch, ctx, X[:2048], G[:2048])
    _, dso = oracle.adjoint(ii, jj, S, X[:2048], G[:2048])
    say("  ds err (2048 points)", normwise_rel_err(ds2.reshape(ii.shape[0], 32), dso))
    assert normwise_rel_err(ds2.reshape(ii.shape[0], 32), dso) <= 1e-12
    h = 1e-4
    for p in sub[:4]:
        for d in (0, 7, 19):
            xp, xm = X[p:p + 1].copy(), X[p:p + 1].copy()
            xp[0, d] += h
            xm[0, d] -= h
            fd = ((oracle.forward(ii, jj, S, xp) - oracle.forward(ii, jj, S, xm)) * G[p]).sum() / (2 * h)
            assert abs(dx[p, d] - fd) <= 1e-7 * max(1.0, abs(fd))
    ctx.close()


def test_adjoint_padded_dx_rows(torch_mod):
    """dx with a row stride above dim: the padding columns keep what the
    caller put there (NaN here), the dim columns match the oracle."""
    torch = torch_mod
    ii, jj, S, X, G = grid_case(7, 2, 5, 0.4, 31, 1500, 24, 150, 5)
    ctx = hddm_lb.Context(0)
    ctx.upload_state(0, ii, jj, S)
    x = torch.from_numpy(X).cuda()
    g = torch.from_numpy(G).cuda()
    dxp = torch.full((X.shape[0], 11), np.nan, dtype=torch.float64, device="cuda")
    ctx.adjoint(x, g, dxp, None, None, hddm_lb.FP64)
    torch.cuda.synchronize()
    dx = dxp.cpu().numpy()
    dxo, _ = oracle.adjoint(ii, jj, S, X, G)
    assert np.isnan(dx[:, 7:]).all()
    assert normwise_rel_err(dx[:, :7], dxo) <= 1e-10
    ctx.close()


@pytest.mark.parametrize("dim,dofs,form,prec", [(20, 32, 0, hddm_lb.FP64), (40, 96, 2, hddm_lb.F32),
                                                (40, 96, 2, hddm_lb.FP64)])
def test_adjoint_dx_bitwise_repeatable(torch_mod, dim, dofs, form, prec):
    """Node segments split every state here; the per-segment dx partials are
    summed in a fixed order, so two runs agree to the last bit. On the
    wave-specialized kernel (form 2) the points are put in a stable order and
    each item's chunks alternate between the two dx waves from the item's
    start, so its two waves' sum is the same in every run too."""
    torch = torch_mod
    ii, jj, S, X, G = grid_case(dim, 2, 5, 0.4, 55, 20000, dofs, 4096, 5)
    ctx = adj_ctx(form)
    ctx.set_option(hddm_lb.OPT_SEGMENT_NODES, 2000)
    ctx.upload_state(0, ii, jj, S)
    a, _ = run_adjoint(torch, ctx, X, G, prec=prec)
    b, _ = run_adjoint(torch, ctx, X, G, prec=prec)
    assert ctx.last_adjoint_config()["kernels"] == (2 if form == 2 else 1)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
    ctx.close()


@pytest.mark.parametrize("prec", [hddm_lb.FP64, hddm_lb.F32])
def test_adjoint_d59_four_states(torch_mod, prec):
    """High-dimensional multistate shape: 4 states, D=59, regular grids
    sum(l-1) <= 3 (281,077 nodes each), dofs=128, 4,096 points in
    [0.01, 0.99]^59 kept 1e-3 off every kink, gbar ~ N(0, 1). dx against the
    oracle on 8 points per state and against central differences on a few
    coordinates; dS against the oracle's B^T g over a 256-point slice."""
    torch = torch_mod
    ctx = hddm_lb.Context(0)
    states = []
    for s in range(4):
        ii, jj, lv, _ = hddm_lb.generate_grid(59, 3, 4, 0.0, 500 + s, 0)
        S = hddm_lb.synthetic_surplus(lv, 128, 600 + s)
        ctx.upload_state(s, ii, jj, S)
        states.append((ii, jj, S))
    say("adjoint D=59 x 4 states, nno", states[0][0].shape[0])
    n = 4096
    X = hddm_lb.synthetic_points(n, 59, 92, 0.01, 0.99, kink_level=4, kink_gap=1e-3)
    sid = np.random.default_rng(93).integers(0, 4, size=n).astype(np.int32)
    G = np.random.default_rng(94).standard_normal((n, 128))
    dx, _ = run_adjoint(torch, ctx, X, G, sid, prec)
    assert np.isfinite(dx).all()
    tdx, tds = TOL_ADJ[prec]
    fd_tol = 1e-7 if prec == hddm_lb.FP64 else 1e-5
    rng = np.random.default_rng(95)
    for s, (ii, jj, S) in enumerate(states):
        pick = rng.choice(np.flatnonzero(sid == s), size=8, replace=False)
        dxo, _ = oracle.adjoint(ii, jj, S, X[pick], G[pick])
        err = normwise_rel_err(dx[pick], dxo)
        say("  state", s, "dx err (8 points)", err)
        assert err <= tdx
        if s == 0:
            h = 1e-4
            for p in pick[:2]:
                for d in (0, 30, 58):
                    xp, xm = X[p:p + 1].copy(), X[p:p + 1].copy()
                    xp[0, d] += h
                    xm[0, d] -= h
                    fd = ((oracle.forward(ii, jj, S, xp) - oracle.forward(ii, jj, S, xm)) * G[p]).sum() / (2 * h)
                    assert abs(dx[p, d] - fd) <= fd_tol * max(1.0, abs(fd))
    m = 256
    _, ds = run_adjoint(torch, ctx, X[:m], G[:m], sid[:m], prec)
    for s, (ii, jj, S) in enumerate(states):
        sel = sid[:m] == s
        _, dso = oracle.adjoint(ii, jj, S, X[:m][sel], G[:m][sel], nthreads=8)
        off, _ = ctx.surplus_offset(s)
        dss = ds[off:off + ii.shape[0] * 128].reshape(ii.shape[0], 128)
        err = normwise_rel_err(dss, dso)
        say("  state", s, "ds err (%d points)" % sel.sum(), err)
        assert err <= tds
    ctx.close()
