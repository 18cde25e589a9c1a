"""Pins the oracle's Lawson-Hanson NNLS (src/NNLS.jl, src/lsqnonneg.jl:30-164).

* scipy.optimize.nnls golden vectors (tests/golden/nnls_cases.npz) for plain and Tikhonov-augmented problems;
* brute force over every active set for n <= 8 (the unique minimizer of a strictly convex problem);
* KKT conditions and the dual identity of the reference's test (test/nnls.jl:84-118);
* Tikhonov: gradient of the active set equals -mu^2 x (test/nnls.jl:255).
"""
from itertools import combinations
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"
SIZES = [(1, 1), (2, 2), (5, 3), (3, 5), (8, 8), (13, 5), (16, 16), (25, 32), (32, 25), (48, 40)]


def rand_nnls(rng, m, n):
    x = rng.random(n)
    x[rng.integers(0, 2)::2] *= -1
    A = rng.random((m, n))
    return A, A @ x


def _unpack(d, key, i):
    off = d[key + "_off"]
    return d[key][off[i]:off[i + 1]]


def test_golden_scipy(oracle):
    d = np.load(GOLD / "nnls_cases.npz")
    for i, (m, n) in enumerate(d["sizes"]):
        A = _unpack(d, "A", i).reshape(m, n)
        b = _unpack(d, "b", i)
        x = _unpack(d, "x", i)
        r = oracle.nnls(A, b)
        f_or = np.sum((A @ r["x"] - b) ** 2)
        f_sc = np.sum((A @ x - b) ** 2)
        assert r["mode"] == 0
        assert f_or <= f_sc * (1 + 1e-9) + 1e-24, (m, n)
        if m >= n:  # full column rank with probability one: unique minimizer
            assert np.allclose(r["x"], x, rtol=1e-8, atol=1e-10), (m, n)
        mu = d["mu"][i]
        xt = _unpack(d, "xt", i)
        rt = oracle.nnls_tikh(A, b, mu)
        assert np.allclose(rt["x"], xt, rtol=1e-7, atol=1e-9), (m, n, mu)


@pytest.mark.parametrize("m,n", [(4, 3), (6, 5), (8, 8), (12, 7)])
def test_brute_force_active_sets(oracle, m, n):
    rng = np.random.default_rng(m * 31 + n)
    for _ in range(5):
        A, b = rand_nnls(rng, m, n)
        best, bx = np.inf, None
        for k in range(n + 1):
            for P in combinations(range(n), k):
                x = np.zeros(n)
                if k:
                    xp, *_ = np.linalg.lstsq(A[:, P], b, rcond=None)
                    if np.any(xp <= 0):
                        continue
                    x[list(P)] = xp
                f = np.sum((A @ x - b) ** 2)
                if f < best:
                    best, bx = f, x
        r = oracle.nnls(A, b)
        assert np.allclose(r["x"], bx, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("m,n", SIZES)
def test_kkt(oracle, m, n):
    rng = np.random.default_rng(7 * m + n)
    A, b = rand_nnls(rng, m, n)
    r = oracle.nnls(A, b)
    x = r["x"]
    w = A.T @ (b - A @ x)
    P = x > 0
    scale = np.linalg.norm(A.T @ b)
    assert np.all(x >= 0)
    assert np.all(w[~P] <= 1e-10 * scale)
    assert np.all(np.abs(w[P]) <= 1e-9 * scale)
    assert r["nsetp"] == P.sum()
    assert sorted(r["idx"][: r["nsetp"]] - 1) == sorted(np.flatnonzero(P))
    assert abs(r["rnorm"] - np.linalg.norm(A @ x - b)) <= 1e-10 * max(1, np.linalg.norm(b))


@pytest.mark.parametrize("m,n", SIZES)
@pytest.mark.parametrize("mu", [1e-6, 1e-2, 10.0])
def test_tikhonov_gradient(oracle, m, n, mu):
    rng = np.random.default_rng(11 * m + n)
    A, b = rand_nnls(rng, m, n)
    r = oracle.nnls_tikh(A, b, mu)
    x = r["x"]
    P = x > 0
    g = A[:, P].T @ (A[:, P] @ x[P] - b)
    assert np.allclose(g, -mu ** 2 * x[P], rtol=1e-7, atol=1e-11 * np.linalg.norm(A.T @ b))
    # rnorm of the augmented system: ||Ax-b||^2 + mu^2 ||x||^2
    assert abs(r["rnorm"] ** 2 - (np.sum((A @ x - b) ** 2) + mu ** 2 * x @ x)) <= 1e-9 * (b @ b)


def test_warm_start_pivot_rule(oracle):
    # w[n] = all(w <= 0) forces the last column first only when no other dual is positive
    A = np.array([[1.0, 0.1], [0.0, 1.0]])
    b = np.array([1.0, 0.0])
    r = oracle.nnls(A, b)
    assert np.allclose(r["x"], [1.0, 0.0])
