"""Pins the CPU oracle (oracle/oracle.c) independently of itself.

Checks: the reference's known-answer test (LinearBasis/tests/InterpolateArray/
test.cpp:24-215 with surplus.plt), the closed form that fixture reduces to,
the hat function's closed form and special cases, exact rational brute force
on tiny grids, central finite differences and exact B^T g for the adjoint.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from conftest import tiny_grid


def test_phi_closed_form_and_special_cases():
    # level 1: i = 0, j = 0 -> 1 everywhere
    for x in (0.0, 0.3, 0.5, 1.0):
        assert oracle.phi(x, 0, 0) == 1.0
    # level 2 boundary nodes: x = 0 and x = 1, support edges at 1/2
    assert oracle.phi(0.0, 2, 0) == 1.0 and oracle.phi(1.0, 2, 2) == 1.0
    assert oracle.phi(0.5, 2, 0) == 0.0 and oracle.phi(0.5, 2, 2) == 0.0
    # level 3 node x = 1/4: peak 1, edges 0 and 1/2 give 0, outside negative
    assert oracle.phi(0.25, 4, 1) == 1.0
    assert oracle.phi(0.0, 4, 1) == 0.0 and oracle.phi(0.5, 4, 1) == 0.0
    assert oracle.phi(0.75, 4, 1) < 0.0
    rng = np.random.default_rng(0)
    for _ in range(200):
        x = rng.random()
        l = int(rng.integers(2, 8))
        i = 2 ** (l - 1)
        j = int(rng.integers(0, i + 1))
        assert oracle.phi(x, i, j) == 1.0 - abs(x * i - j)


def test_reference_known_answer(kat):
    st = kat["state"]
    assert (st["dim"], st["nno"], st["dofs"]) == (59, 119, 118)
    x = np.array(kat["x"])
    np.testing.assert_allclose(x, (np.arange(59) + 1) / 118.0, rtol=0, atol=1e-16)
    y = oracle.forward_point(st["index_i"], st["index_j"], st["surplus"], x)
    np.testing.assert_allclose(y, kat["expected"], rtol=0, atol=kat["epsilon"])
    # the printed values carry 6 significant digits
    np.testing.assert_allclose(y, kat["expected"], rtol=1e-5, atol=1e-5)


def test_reference_fixture_closed_form(kat):
    """Every x_d <= 1/2, so only the root and the x=0 boundary nodes are nonzero:
    y = S_root + sum_d (1 - 2 x_d) S_{node (d, i=2, j=0)}."""
    st = kat["state"]
    ii, jj, S = st["index_i"], st["index_j"], st["surplus"]
    x = np.array(kat["x"])
    y = np.zeros(st["dofs"])
    for k in range(st["nno"]):
        act = np.nonzero((ii[k] != 0) | (jj[k] != 0))[0]
        if len(act) == 0:
            y += S[k]
        elif len(act) == 1 and ii[k, act[0]] == 2 and jj[k, act[0]] == 0:
            y += (1 - 2 * x[act[0]]) * S[k]
    yo = oracle.forward_point(ii, jj, S, x)
    np.testing.assert_allclose(yo, y, rtol=1e-13, atol=1e-13 * np.abs(y).max())


def exact_forward(ii, jj, S, xq):
    """Exact rational B_k(x) with x given as Fractions; returns Fractions [dofs]."""
    out = [Fraction(0)] * S.shape[1]
    for k in range(ii.shape[0]):
        b = Fraction(1)
        for d in range(ii.shape[1]):
            p = 1 - abs(xq[d] * int(ii[k, d]) - int(jj[k, d]))
            if p <= 0:
                b = Fraction(0)
                break
            b *= p
        if b:
            for c in range(S.shape[1]):
                out[c] += b * Fraction(S[k, c])
    return out


@pytest.mark.parametrize("dim,level,seed", [(1, 5, 1), (2, 4, 2), (3, 4, 3), (4, 3, 4)])
def test_forward_exact_rational_brute_force(dim, level, seed):
    ii, jj, lv, ix = tiny_grid(dim, level, seed, p=0.6)
    rng = np.random.default_rng(seed)
    S = rng.uniform(-1, 1, size=(ii.shape[0], 3))
    for _ in range(6):
        xq = [Fraction(int(rng.integers(0, 1 << 20)), 1 << 20) for _ in range(dim)]
        x = np.array([float(v) for v in xq])
        ex = exact_forward(ii, jj, S, xq)
        y = oracle.forward_point(ii, jj, S, x)
        scale = max(1e-300, max(abs(float(v)) for v in ex))
        for c in range(3):
            assert abs(y[c] - float(ex[c])) <= 1e-14 * max(1.0, scale)


def test_batch_and_multistate_match_point():
    ii, jj, lv, ix = tiny_grid(5, 4, 7, p=0.5)
    rng = np.random.default_rng(1)
    S = rng.uniform(-1, 1, size=(ii.shape[0], 5))
    X = rng.random((40, 5))
    Y = oracle.forward(ii, jj, S, X, nthreads=4)
    for p in range(40):
        assert np.array_equal(Y[p], oracle.forward_point(ii, jj, S, X[p]))
    ii2, jj2, lv2, ix2 = tiny_grid(5, 3, 9, p=0.8)
    S2 = rng.uniform(-1, 1, size=(ii2.shape[0], 2))
    ys = oracle.forward_multistate([(ii, jj, S), (ii2, jj2, S2)], [X[0], X[1]])
    assert np.array_equal(ys[0], oracle.forward_point(ii, jj, S, X[0]))
    assert np.array_equal(ys[1], oracle.forward_point(ii2, jj2, S2, X[1]))


def test_adjoint_dx_central_differences():
    dim, L = 4, 4
    ii, jj, lv, ix = tiny_grid(dim, L, 11, p=0.7)
    rng = np.random.default_rng(3)
    S = rng.uniform(-1, 1, size=(ii.shape[0], 3))
    import hddm_lb
    X = hddm_lb.synthetic_points(20, dim, 5, lo=0.01, hi=0.99, kink_level=L)
    G = rng.standard_normal((20, 3))
    dx, ds = oracle.adjoint(ii, jj, S, X, G, nthreads=2)
    h = 1e-4
    for p in range(20):
        for d in range(dim):
            xp, xm = X[p].copy(), X[p].copy()
            xp[d] += h
            xm[d] -= h
            fd = (oracle.forward_point(ii, jj, S, xp) - oracle.forward_point(ii, jj, S, xm)) @ G[p] / (2 * h)
            assert abs(fd - dx[p, d]) <= 1e-8 * max(1.0, abs(fd))


def test_adjoint_dsurplus_exact():
    dim = 3
    ii, jj, lv, ix = tiny_grid(dim, 4, 13, p=0.7)
    rng = np.random.default_rng(4)
    S = rng.uniform(-1, 1, size=(ii.shape[0], 2))
    X = rng.random((10, dim))
    G = rng.standard_normal((10, 2))
    dx, ds = oracle.adjoint(ii, jj, S, X, G, nthreads=3)
    # B^T g with B from the exact rational basis (one-hot surplus columns)
    B = np.zeros((10, ii.shape[0]))
    for p in range(10):
        xq = [Fraction(v) for v in X[p]]
        for k in range(ii.shape[0]):
            b = Fraction(1)
            for d in range(dim):
                q = 1 - abs(xq[d] * int(ii[k, d]) - int(jj[k, d]))
                if q <= 0:
                    b = Fraction(0)
                    break
                b *= q
            B[p, k] = float(b)
    np.testing.assert_allclose(ds, B.T @ G, rtol=1e-13, atol=1e-13)


def test_x86_port_matches_gold():
    import hddm_lb
    ii, jj, lv, ix = hddm_lb.generate_grid(6, 2, 5, 0.4, 21, 3000)
    S = hddm_lb.synthetic_surplus(lv, 12, 2)
    X = hddm_lb.synthetic_points(300, 6, 3)
    st = oracle.X86State(ii, jj, S)
    assert st.nfreqs == int(((ii != 0) | (jj != 0)).sum(axis=1).max())
    y86 = st.forward(X, nthreads=4)
    yg = oracle.forward(ii, jj, S, X, nthreads=4)
    np.testing.assert_allclose(y86, yg, rtol=1e-12, atol=1e-12 * np.abs(yg).max())
