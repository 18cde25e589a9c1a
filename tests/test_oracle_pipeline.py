"""Pins the oracle's pipeline (src/T2mapSEcorr.jl, src/T2partSEcorr.jl): recovery on noiseless data,
save_results! closed forms, threshold semantics and T2part window arithmetic on hand-built distributions.
"""
import math

import numpy as np


def _epg_image(oracle, alphas, T2s, fracs, m=32, TE=10e-3):
    """Noiseless multi-compartment EPG signals, one voxel per row of alphas."""
    X = len(alphas)
    img = np.zeros((X, 1, 1, m), order="F")
    for v, a in enumerate(alphas):
        img[v, 0, 0, :] = sum(f * oracle.epg_const(m, a, TE, t, 1.0) for f, t in zip(fracs, T2s))
    return img


def test_flip_angle_recovered_on_grid(oracle):
    K = 32
    grid = oracle.linrange(50.0, 180.0, K)
    alphas = [grid[20], grid[26], grid[29]]
    img = _epg_image(oracle, alphas, [20e-3, 80e-3], [0.2, 0.8])
    o = oracle.make_t2map_opts(img.shape[:3], 32, 10e-3, 40, (10e-3, 2.0), "none", nRefAngles=K)
    maps, dist, nv = oracle.t2map(img, o)
    assert nv == 3
    assert np.allclose(maps["alpha"][:, 0, 0], alphas, atol=0.05)


def test_save_results_closed_forms(oracle):
    rng = np.random.default_rng(0)
    img = _epg_image(oracle, [150.0, 165.0, 170.0, 140.0], [20e-3, 70e-3, 1.5], [0.15, 0.8, 0.05])
    img = np.asfortranarray(img * (1 + 0.003 * rng.standard_normal(img.shape)) * 1000.0)
    o = oracle.make_t2map_opts(img.shape[:3], 32, 10e-3, 40, (10e-3, 2.0), "chi2", nRefAngles=32, Chi2Factor=1.02)
    maps, dist, nv = oracle.t2map(img, o)
    T2 = oracle.logrange(10e-3, 2.0, 40)
    lg = np.log(T2)
    for v in range(4):
        x = dist[v, 0, 0]
        s = img[v, 0, 0]
        fit = maps["decaycurve"][v, 0, 0]
        res = fit - s
        assert np.isclose(maps["gdn"][v, 0, 0], x.sum(), rtol=1e-12)
        lgm = (x @ lg) / x.sum()
        assert np.isclose(maps["ggm"][v, 0, 0], np.exp(lgm), rtol=1e-12)
        assert np.isclose(maps["gva"][v, 0, 0], np.expm1(((lg - lgm) ** 2 @ x) / x.sum()), rtol=1e-10)
        assert np.isclose(maps["fnr"][v, 0, 0], x.sum() / np.sqrt(np.sum(res ** 2) / 31), rtol=1e-9)
        assert np.isclose(maps["snr"][v, 0, 0], s.max() / np.std(res, ddof=1), rtol=1e-9)
        assert np.isclose(maps["resnorm"][v, 0, 0], np.linalg.norm(res), rtol=1e-9)
        # the fit is the basis at the reported alpha times the distribution
        A = np.stack([oracle.epg_const(32, maps["alpha"][v, 0, 0], 10e-3, t, 1.0) for t in T2], axis=1)
        assert np.allclose(fit, A @ x, rtol=1e-10, atol=1e-12 * s.max())
        assert abs(maps["chi2factor"][v, 0, 0] - 1.02) <= 1.02e-3


def test_threshold_semantics(oracle):
    img = _epg_image(oracle, [150.0, 160.0, 170.0], [20e-3, 80e-3], [0.2, 0.8])
    img[1, 0, 0, 0] = 0.0  # first echo at threshold: skipped (strict >)
    o = oracle.make_t2map_opts(img.shape[:3], 32, 10e-3, 40, (10e-3, 2.0), "none", nRefAngles=8)
    maps, dist, nv = oracle.t2map(img, o)
    assert nv == 2 and np.isnan(maps["gdn"][1, 0, 0]) and np.all(np.isnan(dist[1, 0, 0]))
    o.Threshold = -math.inf
    maps, dist, nv = oracle.t2map(img, o)
    assert nv == 3


def test_t2part_closed_forms(oracle):
    n = 40
    T2 = oracle.logrange(10e-3, 2.0, n)
    sp = np.flatnonzero((T2 >= 10e-3) & (T2 <= 40e-3))
    mp = np.flatnonzero((T2 >= 40e-3) & (T2 <= 2.0))
    rng = np.random.default_rng(1)
    dist = np.asfortranarray(rng.uniform(0, 1, (3, 2, 1, n)))
    dist[2, 1, 0, :] = 0.0
    dist[0, 1, 0, 7] = np.nan
    out = oracle.t2part(dist, (10e-3, 2.0), (10e-3, 40e-3), (40e-3, 2.0))
    for (i, j) in [(0, 0), (1, 0), (2, 0), (1, 1)]:
        d = dist[i, j, 0]
        assert np.isclose(out["sfr"][i, j, 0], d[sp].sum() / d.sum(), rtol=1e-13)
        assert np.isclose(out["mfr"][i, j, 0], d[mp].sum() / d.sum(), rtol=1e-13)
        assert np.isclose(out["sgm"][i, j, 0], np.exp(d[sp] @ np.log(T2[sp]) / d[sp].sum()), rtol=1e-13)
        assert np.isclose(out["mgm"][i, j, 0], np.exp(d[mp] @ np.log(T2[mp]) / d[mp].sum()), rtol=1e-13)
    assert all(np.isnan(out[k][2, 1, 0]) for k in out)  # all-zero distribution: NaN (sums not > 0)
    assert all(np.isnan(out[k][0, 1, 0]) for k in out)  # NaN in the distribution: skipped
    # sigmoid weights: normccdf((T2 - SPWin[2]) / sigma) with sigma = delta / (sqrt(2) |erfinv(-0.8)|)
    import scipy.special
    out = oracle.t2part(dist, (10e-3, 2.0), (10e-3, 40e-3), (40e-3, 2.0), Sigmoid=5e-3)
    sigma = 5e-3 / (np.sqrt(2) * abs(scipy.special.erfinv(-0.8)))
    w = scipy.special.erfc((T2 - 40e-3) / sigma / np.sqrt(2)) / 2
    w[w <= np.finfo(float).eps] = 0.0
    d = dist[1, 0, 0]
    assert np.isclose(out["sfr"][1, 0, 0], d @ w / d.sum(), rtol=1e-12)
