"""Host logic without a device: .plt I/O, grid generator, trie layout, C-ABI symbols."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import hddm_lb
import oracle
from conftest import GOLDEN, PKG, ROOT, tiny_grid


# ------------------------------------------------------------------ .plt

def test_plt_fixture_layout(kat):
    st = kat["state"]
    ii, jj = st["index_i"], st["index_j"]
    act = (ii != 0) | (jj != 0)
    assert act.sum() == 118 and act.sum(axis=1).max() == 1
    assert set(zip(ii[act].tolist(), jj[act].tolist())) == {(2, 0), (2, 2)}


def test_plt_roundtrip_all_formats(tmp_path):
    ii, jj, lv, ix = hddm_lb.generate_grid(7, 2, 5, 0.5, 3, 800)
    S = hddm_lb.synthetic_surplus(lv, 9, 4)
    S[::7, 3] = 0.0  # exercise the sparse surplus encoding
    pc = str(tmp_path / "c.plt")
    hddm_lb.write_plt_compressed(pc, ii, jj, S, level=5)
    r = hddm_lb.read_plt(pc, 7)
    assert (r["dim"], r["nno"], r["dofs"], r["level"]) == (7, ii.shape[0], 9, 5)
    assert np.array_equal(r["index_i"], ii) and np.array_equal(r["index_j"], jj) and np.array_equal(r["surplus"], S)
    for binary in (True, False):
        pu = str(tmp_path / ("u%d.plt" % binary))
        hddm_lb.write_plt_uncompressed(pu, lv, ix, S, binary=binary, level=5)
        r = hddm_lb.read_plt(pu, 7)
        assert np.array_equal(r["index_i"], ii) and np.array_equal(r["index_j"], jj)
        assert np.array_equal(r["surplus"], S)


def test_plt_wide_dims_use_u16_columns(tmp_path):
    # dim > 255 switches the index IA/JA width to 2 bytes
    ii, jj, lv, ix = hddm_lb.generate_grid(300, 1, 2, 0.0, 1, 0)
    S = hddm_lb.synthetic_surplus(lv, 2, 1)
    p = str(tmp_path / "w.plt")
    hddm_lb.write_plt_compressed(p, ii, jj, S)
    r = hddm_lb.read_plt(p, 300)
    assert np.array_equal(r["index_i"], ii) and np.array_equal(r["surplus"], S)


def test_plt_errors(tmp_path):
    with pytest.raises(hddm_lb.HddmLbError) as e:
        hddm_lb.read_plt(str(tmp_path / "missing.plt"), 3)
    assert e.value.code == 6
    ii, jj, lv, ix = hddm_lb.generate_grid(3, 2, 3, 0.0, 1, 0)
    S = hddm_lb.synthetic_surplus(lv, 2, 1)
    p = str(tmp_path / "c.plt")
    hddm_lb.write_plt_compressed(p, ii, jj, S)
    with pytest.raises(hddm_lb.HddmLbError) as e:      # nagents mismatch
        hddm_lb.read_plt(p, 4)
    assert e.value.code == 7
    data = open(p, "rb").read()
    t = str(tmp_path / "t.plt")
    open(t, "wb").write(data[:-5])                      # truncated
    with pytest.raises(hddm_lb.HddmLbError):
        hddm_lb.read_plt(t, 3)
    # a JA (dim) entry out of range
    bad = bytearray(data)
    nnz = int.from_bytes(bad[10 + 16 + 5: 10 + 16 + 9], "little")
    ja0 = 10 + 16 + 9 + 4 * nnz + (ii.shape[0] + 1)
    bad[ja0] = 9
    open(t, "wb").write(bytes(bad))
    with pytest.raises(hddm_lb.HddmLbError) as e:
        hddm_lb.read_plt(t, 3)
    assert e.value.code == 7 and "column" in str(e.value)
    # uncompressed with a level 0 entry
    lv2 = lv.copy()
    lv2[1, 0] = 0
    u = str(tmp_path / "u.plt")
    hddm_lb.write_plt_uncompressed(u, lv2, ix, S)
    with pytest.raises(hddm_lb.HddmLbError) as e:
        hddm_lb.read_plt(u, 3)
    assert e.value.code == 7


def test_plt_level_transform():
    """Uncompressed levels map to i = 2 << (l-2), j = index - 1, j = 0 where i == 0."""
    import tempfile
    lv = np.array([[1, 2, 3], [4, 1, 2]], np.int32)
    ix = np.array([[7, 3, 2], [6, 9, 1]], np.int32)
    S = np.ones((2, 1))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "l.plt")
        hddm_lb.write_plt_uncompressed(p, lv, ix, S)
        r = hddm_lb.read_plt(p, 3)
    assert r["index_i"].tolist() == [[0, 2, 4], [8, 0, 2]]
    assert r["index_j"].tolist() == [[0, 2, 1], [5, 0, 0]]


# ------------------------------------------------------------------ grids

def test_regular_grid_counts():
    # D=4, sum(l-1) <= 4: 401 nodes and 896 active factors
    ii, jj, lv, ix = hddm_lb.generate_grid(4, 4, 5, 1.0, 1, 0)
    assert ii.shape == (401, 4)
    assert int(((ii != 0) | (jj != 0)).sum()) == 896
    # closed form: sum over multi-indices of prod m(l), m = 1, 2, 2, 4, 8, ...
    m = {1: 1, 2: 2, 3: 2, 4: 4, 5: 8}
    import itertools
    tot = sum(int(np.prod([m[l] for l in ls])) for ls in itertools.product(range(1, 6), repeat=4)
              if sum(l - 1 for l in ls) <= 4)
    assert tot == 401
    # D=59, sum <= 3 regular: 281,077 nodes
    g = hddm_lb.generate_grid(59, 3, 4, 1.0, 1, 0)
    assert g[0].shape[0] == 281077


def test_no_refinement_keeps_start_grid():
    """p = 0 below the level cap returns the regular start grid: D=4, sum(l-1)
    <= 2 has 1 + 4*2 + 4*2 + 6*4 = 41 nodes, whatever the cap and target."""
    for maxl, target in ((4, 0), (6, 0), (6, 10 ** 6)):
        ii, jj, lv, ix = hddm_lb.generate_grid(4, 2, maxl, 0.0, 41, target)
        assert ii.shape == (41, 4)
        assert int((lv - 1).sum(axis=1).max()) == 2


def test_adaptive_grid_properties():
    a = hddm_lb.generate_grid(16, 2, 7, 0.3, 1, 60000)
    b = hddm_lb.generate_grid(16, 2, 7, 0.3, 1, 60000)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ii, jj, lv, ix = a
    assert 60000 <= ii.shape[0] < 60000 + 64
    assert (lv - 1).sum(axis=1).max() <= 6
    # unique nodes; level/index consistent with (i, j)
    keys = {tuple(r) for r in np.concatenate([ii, jj], axis=1).tolist()}
    assert len(keys) == ii.shape[0]
    assert np.array_equal(np.where(lv > 1, 1 << (lv - 1), 0), ii)
    assert np.array_equal(np.where(lv > 1, ix - 1, 0), jj)
    # valid 1-D nodes: level 2 -> j in {0, 2}; level >= 3 -> odd j
    l2 = lv == 2
    assert set(np.unique(jj[l2]).tolist()) <= {0, 2}
    hi = lv >= 3
    assert np.all(jj[hi] % 2 == 1)
    c = hddm_lb.generate_grid(16, 2, 7, 0.3, 2, 60000)
    assert not np.array_equal(c[0][:5000], a[0][:5000]) or not np.array_equal(c[1], a[1])


# ------------------------------------------------------------------ layout

def walk_layout(L, X, S_sorted):
    """CPU emulation of one wave's walk over the op stream (the sweep kernel's
    algorithm) for the points X [64, dim]: returns y [64, dofs] and the number
    of ops evaluated."""
    n = X.shape[0]
    P = [np.ones(n) for _ in range(16)]
    y = np.zeros((n, L.dofs))
    q, evals = 0, 0
    while q < L.nops:
        t = L.op_depth[q]
        tt = X[:, L.op_dim[q]] * L.op_i[q] + L.op_negj[q]
        p = P[t] * np.maximum(1.0 - np.abs(tt), 0.0)
        P[t + 1] = p
        evals += 1
        if not p.any():
            assert L.skip[q] > q
            q = int(L.skip[q])
            continue
        if L.op_last[q]:
            y += p[:, None] * S_sorted[L.op_node[q]][None, :]
        q += 1
    return y, evals


def walk_layout_static(L, X, S_sorted, q0=0, q1=None):
    """CPU emulation of the scalar-load walk, which selects the parent product
    by the op's precomputed relation (0 previous product, 1 previous parent,
    2 path stack), stores P[t+1] only for ops with children and jumps by the
    relative skip."""
    n = X.shape[0]
    q1 = L.nops if q1 is None else q1
    P = [np.full(n, np.nan) for _ in range(17)]
    P[0] = np.ones(n)
    cur = np.ones(n)
    base_prev = np.ones(n)
    y = np.zeros((n, L.dofs))
    q, evals = q0, 0
    while q < q1:
        t, rel = L.op_depth[q], L.op_rel[q]
        bt = cur if rel == 0 else (base_prev if rel == 1 else P[t])
        assert not np.isnan(bt).any(), "stack read of a slot never written"
        tt = X[:, L.op_dim[q]] * L.op_i[q] + L.op_negj[q]
        p = bt * np.maximum(1.0 - np.abs(tt), 0.0)
        base_prev, cur = bt, p
        if L.op_kids[q]:
            P[t + 1] = p
        evals += 1
        if not p.any():
            q += int(L.op_relskip[q])
            continue
        if L.op_last[q]:
            y += p[:, None] * S_sorted[L.op_node[q]][None, :]
        q += 1
    return y, evals


@pytest.mark.parametrize("dim,start,level,target,seed", [(4, 4, 5, 0, 1), (8, 2, 5, 2000, 2), (16, 2, 7, 6000, 3)])
def test_layout_trie_walk_matches_oracle(dim, start, level, target, seed):
    ii, jj, lv, ix = hddm_lb.generate_grid(dim, start, level, 0.3, seed, target)
    S = hddm_lb.synthetic_surplus(lv, 5, seed)
    L = hddm_lb.Layout(ii, jj, S)
    nno = ii.shape[0]
    assert sorted(L.order.tolist()) == list(range(nno))
    act = (ii != 0) | (jj != 0)
    # factor sequences of the nodes in trie order are sorted; ops cover exactly the unshared factors
    seqs = [tuple((d, int(ii[k, d]), int(jj[k, d])) for d in np.nonzero(act[k])[0]) for k in L.order]
    assert seqs == sorted(seqs)
    for n in range(nno):
        s = 0
        if n > 0:
            while s < min(len(seqs[n]), len(seqs[n - 1])) and seqs[n][s] == seqs[n - 1][s]:
                s += 1
        ops = range(L.opstart[n], L.opstart[n + 1])
        if len(seqs[n]) == 0:
            assert len(ops) == 1 and L.op_i[ops[0]] == 0 and L.op_last[ops[0]]
            continue
        assert [L.op_depth[q] for q in ops] == list(range(s, len(seqs[n])))
        for q, t in zip(ops, range(s, len(seqs[n]))):
            d, i, j = seqs[n][t]
            assert L.op_dim[q] == d and L.op_i[q] == i and L.op_negj[q] == -j and L.op_node[q] == n
            assert L.op_last[q] == (t == len(seqs[n]) - 1)
            # skip: first op of the first node after the run sharing the first t+1 factors
            e = n + 1
            while e < nno and seqs[e][:t + 1] == seqs[n][:t + 1]:
                e += 1
            assert L.skip[q] == L.opstart[e]
    # relative skips, relations to the stream predecessor and children flags
    for q in range(L.nops):
        assert L.op_relskip[q] == L.skip[q] - q
        t = L.op_depth[q]
        if q == 0:
            assert L.op_rel[q] == 2
        else:
            pt = L.op_depth[q - 1]
            assert L.op_rel[q] == (0 if t == pt + 1 else (1 if t == pt else 2))
        assert L.op_kids[q] == (q + 1 < L.nops and L.op_depth[q + 1] == t + 1)
    # walks with pruning reproduce the oracle on a clustered and a random block
    rng = np.random.default_rng(seed)
    X = rng.random((4096, dim))
    key = np.zeros(4096, np.int64)
    for d in range(dim):
        key = key * 2 + (X[:, d] >= 0.5)
    Xs = X[np.argsort(key, kind="stable")]
    for blk in (Xs[:64], X[:64]):
        yw, evals = walk_layout(L, blk, S[L.order])
        yo = oracle.forward(ii, jj, S, blk)
        np.testing.assert_allclose(yw, yo, rtol=1e-12, atol=1e-12 * np.abs(yo).max())
        yl, evals_l = walk_layout_static(L, blk, S[L.order])
        assert evals_l == evals
        np.testing.assert_allclose(yl, yo, rtol=1e-12, atol=1e-12 * np.abs(yo).max())


def walk_adjoint(L, X, G, S_sorted):
    """CPU emulation of one wave of the adjoint sweep: the trie walk with the
    per-edge sums R_e = sum_{k below e} w_k B_k, flushed into dx[dim(e)] with
    phi'/phi when the walk leaves e. Returns dx [64, dim], dS [nno, dofs] in
    trie order."""
    n, dim = X.shape
    P = [np.ones(n) for _ in range(17)]
    Q = [np.zeros(n) for _ in range(16)]
    R = [np.zeros(n) for _ in range(16)]
    dims = [0] * 16
    dx = np.zeros((n, dim))
    ds = np.zeros((L.nno, L.dofs))
    top, q = -1, 0

    def leave(t):
        nonlocal top
        for tt in range(top, t - 1, -1):
            dx[:, dims[tt]] += R[tt] * Q[tt]
            if tt > 0:
                R[tt - 1] += R[tt]
        top = t

    while q < L.nops:
        t = L.op_depth[q]
        leave(t)
        iv = L.op_i[q]
        u = X[:, L.op_dim[q]] * iv + L.op_negj[q]
        phi = 1.0 - np.abs(u)
        p = P[t] * np.maximum(phi, 0.0)
        P[t + 1] = p
        with np.errstate(divide="ignore", invalid="ignore"):
            Q[t] = np.where(phi > 0, np.where(u > 0, -iv, np.where(u < 0, iv, 0.0)) / phi, 0.0)
        R[t] = np.zeros(n)
        dims[t] = L.op_dim[q]
        if not p.any():
            q = int(L.skip[q])
            continue
        if L.op_last[q]:
            k = L.op_node[q]
            w = G @ S_sorted[k]
            R[t] = w * p
            ds[k] += p @ G
        q += 1
    leave(0)
    return dx, ds


@pytest.mark.parametrize("dim,start,level,target,seed", [(3, 3, 4, 0, 4), (6, 2, 5, 1500, 5), (12, 2, 6, 3000, 6)])
def test_adjoint_edge_sums_match_oracle(dim, start, level, target, seed):
    ii, jj, lv, ix = hddm_lb.generate_grid(dim, start, level, 0.4, seed, target)
    S = hddm_lb.synthetic_surplus(lv, 4, seed)
    L = hddm_lb.Layout(ii, jj, S)
    X = hddm_lb.synthetic_points(64, dim, seed, 0.01, 0.99, kink_level=level, kink_gap=1e-3)
    G = np.random.default_rng(seed).standard_normal((64, 4))
    dx, ds_t = walk_adjoint(L, X, G, S[L.order])
    dxo, dso = oracle.adjoint(ii, jj, S, X, G)
    ds = np.zeros_like(ds_t)
    ds[L.order] = ds_t
    np.testing.assert_allclose(dx, dxo, rtol=0, atol=1e-11 * np.abs(dxo).max())
    np.testing.assert_allclose(ds, dso, rtol=0, atol=1e-13 * np.abs(dso).max())


def flat_adjoint_dx(L, X, G, S_sorted):
    """dx as the wave-specialized adjoint's dx waves form it: for every node
    (trie order) its factors from the factor table, w = g . S, and
    w * (product of the other factors) * phi' added to each factor's dim."""
    nno = L.nno
    ft = L.ftab[:nno]
    d = (ft[:, :, 0] >> 1).astype(np.int64)
    iv = (ft[:, :, 1].astype(np.uint64) << np.uint64(32)).view(np.float64)
    nc = (ft[:, :, 2].astype(np.uint64) << np.uint64(32)).view(np.float64)   # -j / i
    on = ft[:, :, 3] != 0
    u = (X[:, d] + nc[None]) * iv[None]                   # [points, nodes, slots]
    phi = np.maximum(1.0 - np.abs(u), 0.0)
    dp = np.where((phi > 0) & (u != 0), np.where(u > 0, -iv[None], iv[None]), 0.0)
    phi = np.where(on[None], phi, 1.0)
    dp = np.where(on[None], dp, 0.0)
    w = G @ S_sorted.T                                     # [points, nodes]
    dx = np.zeros_like(X)
    for f in range(L.fstride):
        other = np.prod(np.delete(phi, f, axis=2), axis=2)
        term = w * other * dp[:, :, f]
        for p in range(X.shape[0]):
            np.add.at(dx[p], d[:, f], term[p])
    return dx, np.prod(phi, axis=2)


@pytest.mark.parametrize("dim,start,level,target,seed", [(3, 3, 4, 0, 14), (7, 2, 5, 1500, 15), (59, 3, 4, 0, 16)])
def test_factor_table_dx_matches_oracle(dim, start, level, target, seed):
    """The factor table holds each node's factors in the op encoding (the
    basis it gives equals the oracle's), and the leave-one-out dx of the dx
    waves equals the oracle's adjoint."""
    ii, jj, lv, _ = hddm_lb.generate_grid(dim, start, level, 0.4, seed, target)
    S = hddm_lb.synthetic_surplus(lv, 4, seed)
    L = hddm_lb.Layout(ii, jj, S)
    nact = ((ii != 0) | (jj != 0)).sum(axis=1).max()
    assert L.fstride == (4 if nact <= 4 else 8 if nact <= 8 else 16)
    assert L.ftab.shape[0] % 32 == 0 and not L.ftab[L.nno:].any()
    n = 16 if dim > 20 else 48
    X = hddm_lb.synthetic_points(n, dim, seed, 0.01, 0.99, kink_level=level, kink_gap=1e-3)
    G = np.random.default_rng(seed).standard_normal((n, 4))
    dx, B = flat_adjoint_dx(L, X, G, S[L.order])
    y = B @ S[L.order]
    yo = oracle.forward(ii, jj, S, X)
    np.testing.assert_allclose(y, yo, rtol=0, atol=1e-13 * np.abs(yo).max())
    dxo, _ = oracle.adjoint(ii, jj, S, X, G)
    np.testing.assert_allclose(dx, dxo, rtol=0, atol=1e-11 * np.abs(dxo).max())


def test_factor_table_fields():
    """Each node's slots (trie order) hold its active factors in ascending
    dim: {2 dim, high word of i, high word of -j / i (the centre), 1}, then
    zero slots; the low words of i and -j / i are zero, and (x - j / i) i
    equals x i - j (one rounding: x i is exact for i a power of two) bit for
    bit, the value the walk forms with fma(x, i, -j)."""
    ii, jj, lv, _ = hddm_lb.generate_grid(5, 2, 6, 0.4, 21, 600)
    S = hddm_lb.synthetic_surplus(lv, 2, 21)
    L = hddm_lb.Layout(ii, jj, S)
    ft = L.ftab
    hi = lambda v: np.uint32(np.float64(v).view(np.uint64) >> np.uint64(32))
    lo = lambda v: np.uint32(np.float64(v).view(np.uint64) & np.uint64(0xffffffff))
    x = np.random.default_rng(22).random(4096)
    for n in range(L.nno):
        k = L.order[n]
        dims = [d for d in range(5) if ii[k, d] != 0 or jj[k, d] != 0]
        for t, d in enumerate(dims):
            i, j = int(ii[k, d]), int(jj[k, d])
            c = -float(j) / i                     # -0.0 for j = 0, as the builders compute it
            assert lo(i) == 0 and lo(c) == 0
            assert tuple(ft[n, t]) == (2 * d, hi(i), hi(c), 1)
        assert not ft[n, len(dims):].any()
    for i in (2, 4, 8, 16, 32):
        for j in range(0, i + 1):
            u_centre = (x + np.float64(-float(j) / i)) * i
            u_fma = x * i - j                     # x * i exact: one rounding
            assert np.array_equal(u_centre.view(np.uint64), u_fma.view(np.uint64))


def test_layout_rejects_bad_input():
    with pytest.raises(hddm_lb.HddmLbError) as e:
        hddm_lb.Layout(np.zeros((0, 3), np.int32), np.zeros((0, 3), np.int32), np.zeros((0, 2)))
    assert e.value.code == 1


# ------------------------------------------------------------------ C ABI

def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(hddm_lb_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions(os.path.join(ROOT, "include", "hddm_lb.h"))
    assert len(names) >= 29
    lib = ctypes.CDLL(hddm_lb.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_paths_fail_cleanly():
    n = hddm_lb.device_count()
    if n > 0:
        pytest.skip("a device is visible")
    with pytest.raises(hddm_lb.HddmLbError) as e:
        hddm_lb.Context(0)
    assert e.value.code == 2 and "no HIP device" in str(e.value)
