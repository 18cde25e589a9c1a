"""Pins the oracle's regularization searches and scalar optimizers (src/lsqnonneg.jl, src/optimization.jl,
src/splines.jl) against properties the mathematics fixes and the reference's own tests check.
"""
from pathlib import Path

import numpy as np
import pytest
import scipy.optimize

GOLD = Path(__file__).resolve().parent / "golden"
SIZES = [(2, 2), (5, 3), (8, 8), (13, 5), (16, 16), (25, 32), (32, 25), (48, 40)]


def rand_nnls(rng, m, n):
    x = rng.random(n)
    x[rng.integers(0, 2)::2] *= -1
    A = rng.random((m, n))
    return A, A @ x


def epg_basis(oracle, m=48, alpha=150.0, n=40, TE=10e-3):
    T2 = oracle.logrange(10e-3, 2.0, n)
    return np.stack([oracle.epg_const(m, alpha, TE, t, 1.0) for t in T2], axis=1), T2


# ------------------------------------------------------------------ scalar optimizers
def test_brent_root(oracle):
    # test/optimization.jl: roots of sin, exp-1 within bracketing intervals
    x, f = oracle.brent_root(np.sin, 2.0, 4.0, xatol=1e-14)
    assert abs(x - np.pi) <= 1e-12
    x, f = oracle.brent_root(lambda t: np.expm1(t) - 0.5, -1.0, 2.0, xatol=1e-14)
    assert abs(x - np.log(1.5)) <= 1e-12
    x, f = oracle.brent_root(lambda t: t ** 3 - 2, 0.0, 2.0, ftol=1e-6)
    assert abs(f) <= 1e-6


def test_bracket_root_monotonic(oracle):
    a, b, fa, fb = oracle.bracket_root_monotonic(lambda t: t - 10.3, -4.0, 1.0, dilate=1.5, mono=1.0, maxiters=20)
    assert a <= 10.3 <= b and fa < 0 < fb
    a, b, fa, fb = oracle.bracket_root_monotonic(lambda t: 2.0 - t, 5.0, 1.0, dilate=1.5, mono=-1.0, maxiters=20)
    assert a <= 2.0 <= b


def test_brent_minimize(oracle):
    x, f = oracle.brent_minimize(lambda t: np.sin(t) ** 2 + 0.1 * (t - 3) ** 2, 2.0, 4.5, xrtol=0.0, xatol=1e-10)
    r = scipy.optimize.minimize_scalar(lambda t: np.sin(t) ** 2 + 0.1 * (t - 3) ** 2, bounds=(2.0, 4.5),
                                       method="bounded", options=dict(xatol=1e-12))
    assert abs(x - r.x) <= 1e-7


def test_hermite_minimize_exact_on_cubics(oracle):
    # a cubic is its own Hermite interpolant: the surrogate minimum is the cubic's minimum on [a, b]
    rng = np.random.default_rng(5)
    for _ in range(500):
        c = rng.normal(size=4)
        a, b = sorted(rng.uniform(-2, 2, size=2))
        f = lambda t: ((c[3] * t + c[2]) * t + c[1]) * t + c[0]
        df = lambda t: (3 * c[3] * t + 2 * c[2]) * t + c[1]
        x, u = oracle.hermite_minimize(a, b, f(a), f(b), df(a), df(b))
        grid = np.linspace(a, b, 20001)
        assert u <= f(grid).min() + 1e-9 * max(1, abs(f(grid)).max())
        assert abs(u - f(x)) <= 1e-9 * max(1, abs(f(grid)).max())


def test_lcurve_corner_mock(oracle):
    # test/nnls.jl:349-365: (xi, eta) = (mu, 1/mu) has its maximum curvature at mu = 1
    x = oracle.lcurve_corner(lambda lx: (np.exp(lx), np.exp(-lx)), np.log(0.1), np.log(10.0), 1e-6, 1e-6, 0.0)
    assert abs(x) <= 1e-3


# ------------------------------------------------------------------ singular values / GCV
def test_svdvals_golden(oracle):
    d = np.load(GOLD / "svd_cases.npz")
    for k, (m, n) in enumerate(d["shapes"]):
        M = d["M"][d["M_off"][k]:d["M_off"][k + 1]].reshape(m, n)
        s = d["s"][d["s_off"][k]:d["s_off"][k + 1]]
        assert np.max(np.abs(oracle.svdvals(M) - s)) <= 1e-13 * s[0]


@pytest.mark.parametrize("m,n", SIZES)
def test_gcv_dof_trace_formula(oracle, m, n):
    # test/nnls.jl:447: dof = tr(I - A (A'A + mu^2 I)^-1 A')
    rng = np.random.default_rng(m + 3 * n)
    A, _ = rand_nnls(rng, m, n)
    mu = np.exp(rng.normal())
    g = oracle.svdvals(A)
    tr = np.trace(np.eye(m) - A @ np.linalg.solve(A.T @ A + mu ** 2 * np.eye(n), A.T))
    assert abs(oracle.gcv_dof(m, n, g, mu) - tr) <= 1e-8 * max(1, tr)


def test_gcv_is_a_local_minimum(oracle):
    rng = np.random.default_rng(2)
    A, _ = epg_basis(oracle)
    xt = np.zeros(40); xt[8] = 0.2; xt[20] = 0.8
    b = A @ xt + 0.003 * rng.standard_normal(48)
    x, mu, chi2 = oracle.lsqnonneg(A, b, "gcv")
    s = np.linalg.svd(A, compute_uv=False)

    def G(logmu):
        m_ = np.exp(logmu)
        xa, _ = scipy.optimize.nnls(np.vstack([A, m_ * np.eye(40)]), np.concatenate([b, np.zeros(40)]), maxiter=4000)
        dof = 8 + np.sum(m_ ** 2 / (s ** 2 + m_ ** 2))
        return np.sum((A @ xa - b) ** 2) / dof ** 2

    lm = np.log(mu)
    g0 = G(lm)
    assert g0 <= G(lm - 0.05) * (1 + 1e-6) and g0 <= G(lm + 0.05) * (1 + 1e-6)
    xa, _ = scipy.optimize.nnls(np.vstack([A, mu * np.eye(40)]), np.concatenate([b, np.zeros(40)]), maxiter=4000)
    assert np.allclose(x, xa, rtol=1e-6, atol=1e-9)


# ------------------------------------------------------------------ chi2 / mdp / lcurve / none
@pytest.mark.parametrize("m,n", SIZES)
def test_chi2_hits_target(oracle, m, n):
    # test/nnls.jl:399-424
    rng = np.random.default_rng(13 * m + n)
    A, b = rand_nnls(rng, m, n)
    xu = oracle.nnls(A, b)
    r2min, r2max = xu["rnorm"] ** 2, b @ b
    target = min(np.sqrt(r2min * r2max) / max(r2min, 1e-300), 1.01 + 0.99 * rng.random())
    x, mu, chi2 = oracle.lsqnonneg(A, b, "chi2", target)
    if r2min <= 1e-12 or np.sum(xu["x"] ** 2) == 0:
        assert mu >= 0
    else:
        assert mu > 0
        assert abs(chi2 - target) <= 1e-3 * target
        assert abs(np.sum((A @ x - b) ** 2) / r2min - chi2) <= 1e-6 * chi2


@pytest.mark.parametrize("m,n", SIZES)
def test_mdp_hits_discrepancy(oracle, m, n):
    # test/nnls.jl:483-519
    rng = np.random.default_rng(17 * m + n)
    A, b = rand_nnls(rng, m, n)
    xu = oracle.nnls(A, b)
    r2min, r2max = xu["rnorm"] ** 2, b @ b
    r2t = r2min + (r2max - r2min) / 10
    x, mu, chi2 = oracle.lsqnonneg(A, b, "mdp", np.sqrt(r2t))
    r2 = np.sum((A @ x - b) ** 2)
    if mu <= 0:
        assert mu == 0 and chi2 == 1
    else:
        assert abs(r2 - r2t) <= 2e-3 * r2t + 1e-12


def test_lcurve_solution_is_tikhonov_at_mu(oracle):
    rng = np.random.default_rng(4)
    A, _ = epg_basis(oracle, m=32, alpha=170.0)
    xt = np.zeros(40); xt[10] = 0.3; xt[25] = 0.7
    b = A @ xt + 0.004 * rng.standard_normal(32)
    x, mu, chi2 = oracle.lsqnonneg(A, b, "lcurve")
    assert np.exp(-8) <= mu <= np.exp(2)
    xa, _ = scipy.optimize.nnls(np.vstack([A, mu * np.eye(40)]), np.concatenate([b, np.zeros(40)]), maxiter=4000)
    assert np.allclose(x, xa, rtol=1e-6, atol=1e-9)
    xu = oracle.nnls(A, b)
    assert abs(chi2 - np.sum((A @ x - b) ** 2) / xu["rnorm"] ** 2) <= 1e-8 * chi2


def test_none_is_plain_nnls(oracle):
    rng = np.random.default_rng(9)
    A, b = rand_nnls(rng, 20, 10)
    x, mu, chi2 = oracle.lsqnonneg(A, b, "none")
    assert mu == 0 and chi2 == 1
    assert np.allclose(x, oracle.nnls(A, b)["x"])
