"""Pins the oracle's EPG recursions (src/EPGdecaycurve.jl) without retyping them.

* closed form: alpha = beta = 180 gives dc[i] = exp(-i TE / T2) for any T1;
* golden vectors from an independent full-state complex EPG simulation (tests/golden/make_golden.py);
* cross-form agreement, as the reference's own test (test/epg.jl:36-121): the split DualFlat / DualVector
  recursions and the complex rotation-matrix form agree to rtol sqrt(eps), atol 10 eps;
* d/dalpha against central finite differences with the reference's O(h^2) bars (test/epg.jl:161-172).
"""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("etl", [4, 5, 17, 48, 64])
@pytest.mark.parametrize("T1", [0.1, 1.0, 151.0])
def test_closed_form_180(oracle, etl, T1):
    TE, T2 = 10e-3, 0.05
    ref = np.exp(-np.arange(1, etl + 1) * TE / T2)
    assert np.allclose(oracle.epg_const(etl, 180.0, TE, T2, T1), ref, rtol=1e-14, atol=1e-16)
    assert np.allclose(oracle.epg_general(etl, 180.0, TE, T2, T1, 180.0), ref, rtol=1e-14, atol=1e-16)
    assert np.allclose(oracle.epg_basic(etl, 180.0, TE, T2, T1, 180.0), ref, rtol=1e-14, atol=1e-16)


def test_golden_full_state_simulation(oracle):
    d = np.load(GOLD / "epg_cases.npz")
    for p, c in zip(d["params"], d["curves"]):
        etl = int(p[0])
        a, TE, T2, T1, beta = p[1:]
        ref = c[:etl]
        g = oracle.epg_general(etl, a, TE, T2, T1, beta)
        assert np.allclose(g, ref, rtol=np.sqrt(np.finfo(float).eps), atol=10 * np.finfo(float).eps), p
        if beta == 180.0:
            assert np.allclose(oracle.epg_const(etl, a, TE, T2, T1), ref, rtol=1.5e-8, atol=2.3e-15), p


@pytest.mark.parametrize("etl", [4, 5, 6, 7, 33, 64])
def test_cross_forms(oracle, etl):
    tol = dict(rtol=np.sqrt(np.finfo(float).eps), atol=10 * np.finfo(float).eps)
    TE, T2, T1 = 39e-3, 1.1, 151.0
    for a in (165.0, 120.0, 95.0):
        c = oracle.epg_const(etl, a, TE, T2, T1)
        assert np.allclose(c, oracle.epg_general(etl, a, TE, T2, T1, 180.0), **tol)
        assert np.allclose(c, oracle.epg_basic(etl, a, TE, T2, T1, 180.0), **tol)
        for beta in (150.0, 172.0):
            assert np.allclose(oracle.epg_general(etl, a, TE, T2, T1, beta), oracle.epg_basic(etl, a, TE, T2, T1, beta),
                               **tol)


@pytest.mark.parametrize("beta", [180.0, 176.0])
def test_dalpha_finite_differences(oracle, beta):
    # test/epg.jl:161-172: J*dx vs f(x+dx)-f(x) at O(dx^2): 5e-4, 5e-8, 5e-12
    ETL, a, TE, T2, T1 = 10, 169.0, 9e-3, 10.1e-3, 0.98
    y, dy = oracle.epg_dalpha(ETL, a, TE, T2, T1, beta)
    assert np.allclose(y, oracle.epg_basic(ETL, a, TE, T2, T1, beta), atol=10 * np.finfo(float).eps)
    for h, bar in ((1e-2, 5e-4), (1e-4, 5e-8), (1e-6, 5e-12)):
        d = a * h
        y1 = oracle.epg_basic(ETL, a + d, TE, T2, T1, beta)
        assert np.max(np.abs(dy * d - (y1 - y))) <= bar


def test_sind_cosd_exact_points(oracle):
    for x, v in ((0.0, 0.0), (30.0, 0.5), (90.0, 1.0), (150.0, 0.5), (180.0, 0.0), (270.0, -1.0)):
        assert oracle.lib().orc_sind(x) == v
    assert oracle.lib().orc_cosd(60.0) == 0.5
    assert abs(oracle.lib().orc_sind(45.0) - np.sqrt(0.5)) <= 1.2e-16


def test_erfinv_matches_scipy(oracle):
    import scipy.special

    for y in (-0.999, -0.8, -0.3, 1e-9, 0.5, 0.95):
        assert abs(oracle.lib().orc_erfinv(y) - scipy.special.erfinv(y)) <= 4e-16 * max(1, abs(scipy.special.erfinv(y)))
