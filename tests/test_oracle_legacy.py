"""The oracle's legacy path (legacy = true) pinned by FITPACK itself, closed forms and Julia's documented ranges.

- Dierckx.Spline1D(x, y; k, s = 0) (src/splines.jl:312-315) wraps FITPACK's curfit; scipy.interpolate.splrep/splev
  wrap the same FITPACK routines, so knots, coefficients and values must agree bit for bit (a wrong knot rule for
  even degree, a swapped Givens sign or a dropped back-substitution term fails at once);
- splines reproduce polynomials of their degree, so a cubic sampled on the 0.001 grid has a known argmin;
- Julia's range start:step:stop (the sample grid of spline_opt_legacy / spline_root_legacy, src/splines.jl:419-446)
  has documented lengths and exact decimal elements (0.0:0.1:0.3 ends at 0.3, not 3 * 0.1); on the rational path
  every element is the correctly rounded quotient k / 1000, and none of the grid points used lies within 2^-90 of a
  rounding midpoint, where Julia's twice-precision arithmetic could round otherwise;
- chi2 legacy (src/lsqnonneg.jl:521-532, 595-636): mu doubles from 1e-3 until res2 >= chi2 res2_min; the result is
  the grid point of 0:0.001:mu_last nearest the target on scipy's spline through the (mu, res2) pairs, and x is the
  Tikhonov solution at that mu.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.interpolate as si


@pytest.mark.parametrize("k", [1, 2, 3])
@pytest.mark.parametrize("m", [2, 3, 4, 5, 8, 21, 64])
def test_spline_matches_fitpack_bitwise(oracle, k, m):
    if m <= k:
        pytest.skip("needs m > k")
    rng = np.random.default_rng(100 * k + m)
    x = np.sort(rng.uniform(-3.0, 7.0, m))
    x = np.unique(x)
    y = rng.normal(size=len(x)) * 10.0 ** rng.uniform(-3, 3)
    t, c = oracle.spline_interp(x, y, k)
    ts, cs, _ = si.splrep(x, y, k=k, s=0)
    assert np.array_equal(t, ts)
    assert np.array_equal(c, cs[:len(x)])
    xx = np.concatenate([x, np.linspace(x[0], x[-1], 101)])
    assert np.array_equal(oracle.spline_eval(t, c, k, xx), si.splev(xx, (ts, cs, k)))


def test_spline_reproduces_cubics(oracle):
    x = np.linspace(50.0, 180.0, 8)
    p = np.polynomial.Polynomial([3.0, -0.2, 1.5e-3, -4e-6])
    t, c = oracle.spline_interp(x, p(x), 3)
    xx = np.linspace(50.0, 180.0, 513)
    assert np.allclose(oracle.spline_eval(t, c, 3, xx), p(xx), rtol=0, atol=1e-12)


def test_julia_ranges(oracle):
    r, n = oracle.jlrange(0.0, 0.1, 0.3)
    assert n == 4 and r[-1] == 0.3 and r[-1] != 3 * 0.1
    r, n = oracle.jlrange(1.0, 0.1, 2.0)
    assert n == 11 and r[3] == 1.3 and r[-1] == 2.0
    r, n = oracle.jlrange(50.0, 0.001, 180.0)
    assert n == 130001 and r[0] == 50.0 and r[-1] == 180.0
    k = 50000 + np.arange(n)
    assert np.array_equal(r, k / 1000.0)  # IEEE division of exact integers: the correctly rounded quotient
    for p in range(0, 12):
        r, n = oracle.jlrange(0.0, 0.001, 0.001 * 2.0 ** p)
        assert n == 2 ** p + 1 and r[-1] == 0.001 * 2.0 ** p
        assert np.array_equal(r, np.arange(n) / 1000.0)
    r, n = oracle.jlrange(0.0, 0.001, 0.0)
    assert n == 1 and r[0] == 0.0


def _midpoint_gap(k):
    """Relative distance of k/1000 from the nearest rounding midpoint of binary64."""
    q = Fraction(k, 1000)
    f = float(q)
    e = math.frexp(f)[1] - 53
    ulp = Fraction(2) ** e
    mids = [Fraction(f) + ulp / 2, Fraction(f) - ulp / 2]
    return min(abs(q - mp) for mp in mids) / q


def test_grid_points_are_far_from_rounding_midpoints():
    worst = min(_midpoint_gap(k) for k in range(50000, 180001))
    worst = min(worst, min(_midpoint_gap(k) for k in range(1, 1 << 14)))
    assert worst > Fraction(1, 2 ** 90)


def test_spline_opt_legacy_on_a_cubic(oracle):
    # loss with its minimum at 120.0004 on [50, 180]: the spline through 8 samples is the cubic itself, so the best
    # sample of 50:0.001:180 is 120.000 (the cubic rises by ~1e-13 per 1e-4 there, far above rounding)
    a = 120.0004
    p = np.polynomial.Polynomial([0.0, 0.0, 1.0]) * np.polynomial.Polynomial([1.0, 1e-3])  # (1 + 1e-3 s) s^2
    f = lambda al: p(al - a)
    x = np.linspace(50.0, 180.0, 8)
    xo, yo = oracle.spline_opt_legacy(x, f(x))
    assert xo == 120.0
    assert yo == pytest.approx(f(120.0), abs=1e-12)


def test_spline_root_legacy_linear(oracle):
    # two points: degree 1, the sample of 0:0.001:0.001 nearer the target; three or more: the nearest grid point to
    # the root of the (here exact) spline
    assert oracle.spline_root_legacy([0.0, 0.001], [1.0, 2.0], 1.4) == 0.0
    assert oracle.spline_root_legacy([0.0, 0.001], [1.0, 2.0], 1.6) == 0.001
    x = np.array([0.0, 0.001, 0.002, 0.004, 0.008])
    y = 2.0 + 5.0 * x
    assert oracle.spline_root_legacy(x, y, 2.0 + 5.0 * 0.0063) == 0.006


def test_surrogate_search_legacy_evaluates_every_angle(oracle):
    alphas = np.linspace(50.0, 180.0, 8)
    a = 151.2346
    x, trace = oracle.surrogate_search(alphas, 8, lambda al: ((al - a) ** 2 * (1 + 1e-3 * (al - a)), 0.0),
                                       legacy=True)
    assert sorted(trace) == list(range(1, 9))
    assert x == 151.235


def _synthetic_voxel(seed, m=32, n=24):
    rng = np.random.default_rng(seed)
    t = 10e-3 * np.arange(1, m + 1)
    T2 = np.logspace(np.log10(10e-3), np.log10(2.0), n)
    A = np.exp(-t[:, None] / T2[None, :])
    x = np.zeros(n)
    x[rng.integers(2, 8)] = 0.3
    x[rng.integers(10, 20)] = 0.7
    b = A @ x
    b = b + rng.normal(scale=2e-3, size=m)
    return A / b.max(), b / b.max()


@pytest.mark.parametrize("seed", range(6))
def test_chi2_legacy_against_fitpack(oracle, seed):
    A, b = _synthetic_voxel(seed)
    chi2f = 1.02
    x, mu, chi2 = oracle.lsqnonneg(A, b, "chi2_legacy", chi2f)
    x0 = oracle.nnls(A, b)
    r2min = x0["rnorm"] ** 2
    target = chi2f * r2min

    def r2(mu):
        s = oracle.nnls_tikh(A, b, mu)
        return float(np.sum((A @ s["x"] - b) ** 2)), s["x"]

    mus, r2s = [0.0], [r2min]
    mun = 1e-3
    while True:
        v, _ = r2(mun)
        mus.append(mun)
        r2s.append(v)
        if v >= target:
            break
        mun *= 2
    k = min(3, len(mus) - 1)
    tck = si.splrep(mus, r2s, k=k, s=0)
    grid = np.arange(int(round(mus[-1] * 1000)) + 1) / 1000.0
    expect = grid[np.argmin(np.abs(si.splev(grid, tck) - target))]
    assert mu == expect
    assert mu * 1000 == round(mu * 1000)
    if mu > 0:
        v, xs = r2(mu)
        assert np.allclose(x, xs, rtol=1e-12, atol=1e-15)
        assert chi2 == pytest.approx(v / r2min, rel=1e-10)
    else:
        assert np.allclose(x, x0["x"], rtol=0, atol=0) and chi2 == 1.0
