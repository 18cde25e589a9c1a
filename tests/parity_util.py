"""Per-voxel parity against the oracle with a rounding-stability classification.

The regularization searches of the reference end on tolerance tests (Brent's |x - xm| <= tol2 - (b - a)/2 in
src/optimization.jl:319-413, the L-curve's xtol/Ptol/Ctol tests and isapprox-keyed caches in
src/lsqnonneg.jl:872-972). A voxel whose search reaches such a test with a margin at the rounding level takes one
more or one fewer step after a change in the last bit of its input, and its mu moves by up to the search tolerance
(xatol = 1e-4 in log mu for gcv), which moves the distribution by ~1e-4. The oracle shows this on itself: the same
voxels change by ~1e-4 when the image is scaled by 1 + j 2^-52.

So every voxel is classified by the oracle alone: it is run on the image and on 16 copies scaled by 1 + j 2^-52,
j = +-1..+-8 (each echo moves by at most a few ulps; the max-normalized signal by ulps), and with its NNLS dot
products summed in two other orders (reversed, and as two interleaved partial sums), on the image and on the copies
scaled by 1 +- 2^-52. The summation orders matter where the regularized problem is ill-conditioned (few echoes,
small mu): there the device's own summation order moves the result more than an input ulp does. A voxel is stable
when every member agrees with the unperturbed oracle to a tenth of the bar, and, for the L-curve, when its final
choice is not a near-tie: the corner is the cached point of largest Menger curvature, and near convergence the
neighbouring points' curvatures differ by as little as 1e-7 relative, below the rounding noise of curvatures formed
from nearby log-residuals (rn^2 - mu^2 |x|^2 cancels); a voxel whose best and second-best curvature are within
1e-6 relative is a tie, its answer one of the tied points (distribution relative L2 and MWF <= 1e-7, alpha
<= 1e-3 deg). Stable voxels must meet the FP64 bar on every voxel: alpha <= 0.01 deg, distribution relative L2
<= 1e-6, MWF <= 1e-6, and the derived maps to 1e-6 relative. Unstable voxels are listed by id and bounded by
their own rounding spread: the GPU may differ from the oracle no more than 4x the largest change the ensemble
produced (and never more than the mixed-precision bar, MWF <= 1e-3).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

# folder for each case's full per-voxel record (parity_<case>.json); unset: the summary line is printed only
OUT = os.environ.get("DMI_PARITY_DUMPS")

SCALES = tuple(1.0 + j * 2.0 ** -52 for k in range(1, 9) for j in (k, -k))
ORDERED = tuple((order, 1.0 + j * 2.0 ** -52) for order in (1, 2) for j in (0, 1, -1))  # (sum order, scale)
STABLE_DIST, STABLE_MWF, STABLE_ALPHA = 1e-7, 1e-7, 1e-3
LCURVE_TIE = 1e-6        # relative curvature gap of the L-curve's final argmax below which the choice is a tie
TIE_DIST, TIE_MU = 1e-3, 2e-3  # a tie moves mu by a few search steps (xtol 1e-4 in log mu) at most
BAR_DIST, BAR_MWF, BAR_ALPHA = 1e-6, 1e-6, 1e-2
MAPS_REL = 1e-6


def _mwf(oracle, dist, T2R):
    return oracle.t2part(dist, T2R, (T2R[0], 40e-3), (40e-3, T2R[1]))["sfr"]


def _rel_l2(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)


def oracle_ensemble(oracle, img, oo, T2R, *, vox=None, alpha_in=None, nthreads=None):
    """Oracle maps and distribution on img, plus the per-voxel spread over the scaled copies.

    With vox (linear voxel indices), only those voxels are run (t2map_subset). Returns (maps, dist, sel, spread),
    sel the boolean (X,Y,Z) selection of voxels the oracle processed, spread a dict of per-voxel arrays over sel:
    dist (relative L2), mwf, alpha, mu (relative)."""
    nthreads = nthreads or min(os.cpu_count() or 1, 16)

    def run(im, order=0):
        oracle.lib().orc_set_sum_order(order)
        try:
            if vox is None:
                m, d, _ = oracle.t2map(im, oo, alpha_in=alpha_in, nthreads=nthreads)
            else:
                m, d = oracle.t2map_subset(im, oo, vox, nthreads=nthreads)
        finally:
            oracle.lib().orc_set_sum_order(0)
        return m, d

    margin = None
    if int(getattr(oo, "Reg", -1)) == oracle.REG["lcurve"]:
        import ctypes as C

        margin = np.full(int(np.prod(img.shape[:3])), np.inf)
        oracle.lib().orc_set_margin_output(margin.ctypes.data_as(C.POINTER(C.c_double)))
    try:
        maps, dist = run(img)
    finally:
        if margin is not None:
            oracle.lib().orc_set_margin_output(None)
    if vox is None:
        sel = ~np.isnan(maps["gdn"])
    else:
        sel = np.zeros(int(np.prod(img.shape[:3])), bool)
        sel[np.asarray(vox)] = True
        sel = sel.reshape(img.shape[:3], order="F")
    d0, a0, mu0 = dist[sel], maps["alpha"][sel], maps["mu"][sel]
    w0 = _mwf(oracle, dist, T2R)[sel]
    spread = dict(dist=np.zeros(d0.shape[0]), mwf=np.zeros(d0.shape[0]), alpha=np.zeros(d0.shape[0]),
                  mu=np.zeros(d0.shape[0]),
                  tie=(margin.reshape(img.shape[:3], order="F")[sel] < LCURVE_TIE) if margin is not None
                  else np.zeros(d0.shape[0], bool))
    for order, s in [(0, s) for s in SCALES] + list(ORDERED):
        m, d = run(img * s, order)
        spread["dist"] = np.maximum(spread["dist"], _rel_l2(d[sel], d0))
        spread["mwf"] = np.maximum(spread["mwf"], np.abs(_mwf(oracle, d, T2R)[sel] - w0))
        spread["alpha"] = np.maximum(spread["alpha"], np.abs(m["alpha"][sel] - a0))
        mu = m["mu"][sel]
        spread["mu"] = np.maximum(spread["mu"], np.abs(mu - mu0) / np.maximum(np.abs(mu0), 1e-300))
    return maps, dist, sel, spread


def check(name, oracle, T2R, sel, spread, maps, dist, omaps, odist, *, derived=("gdn", "ggm", "gva", "fnr", "snr"),
          extra_check=None):
    """Asserts the per-voxel bar on stable voxels and the spread bound on unstable ones; writes a summary."""
    d_g, d_o = dist[sel], odist[sel]
    rel = _rel_l2(d_g, d_o)
    da = np.abs(maps["alpha"][sel] - omaps["alpha"][sel])
    mwf = np.abs(_mwf(oracle, dist, T2R)[sel] - _mwf(oracle, odist, T2R)[sel])
    tie = spread["tie"]
    stable = ((spread["dist"] <= STABLE_DIST) & (spread["mwf"] <= STABLE_MWF) & (spread["alpha"] <= STABLE_ALPHA)
              & ~tie)
    lin = np.ravel_multi_index(np.nonzero(sel), sel.shape, order="F")  # column-major ids, in dist[sel] order
    unstable_ids = lin[~stable]

    def q(x):
        x = np.asarray(x, float)
        if x.size == 0:
            return None
        return dict(max=float(np.max(x)), p99=float(np.quantile(x, 0.99)), median=float(np.median(x)))

    s = dict(case=name, voxels=int(sel.sum()), stable=int(stable.sum()), unstable_ids=[int(i) for i in unstable_ids],
             lcurve_ties=int(tie.sum()),
             stable_dist_rel_l2=q(rel[stable]), stable_alpha_abs=q(da[stable]), stable_mwf_abs=q(mwf[stable]),
             unstable_dist_rel_l2=[float(x) for x in rel[~stable]],
             unstable_spread_dist=[float(x) for x in spread["dist"][~stable]],
             unstable_mwf_abs=[float(x) for x in mwf[~stable]])
    mu_rel = None
    if "mu" in maps and "mu" in omaps:
        mu_rel = np.abs(maps["mu"][sel] - omaps["mu"][sel]) / np.maximum(np.abs(omaps["mu"][sel]), 1e-300)
        s["stable_mu_rel"] = q(mu_rel[stable])
    bad = lin[stable & ((rel > BAR_DIST) | (da > BAR_ALPHA) | (mwf > BAR_MWF))]
    s["stable_failures"] = [int(i) for i in bad]
    if OUT:
        Path(OUT).mkdir(parents=True, exist_ok=True)
        with open(Path(OUT) / f"parity_{name}.json", "w") as f:
            json.dump(s, f, indent=1)
    print(json.dumps({k: v for k, v in s.items() if not k.startswith("unstable_")} |
                     dict(unstable=len(unstable_ids), unstable_max_dist=float(rel[~stable].max(initial=0.0)))),
          flush=True)
    # alpha is decided before regularization: every voxel
    assert np.max(da, initial=0.0) <= BAR_ALPHA, s
    # every stable voxel meets the FP64 bar
    assert bad.size == 0, s
    if mu_rel is not None:
        assert np.all(mu_rel[stable] <= 1e-5), s
    for k in derived:
        a, b = maps[k][sel][stable], omaps[k][sel][stable]
        err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        assert np.all(err <= MAPS_REL), (k, float(err.max(initial=0.0)))
    # unstable voxels: within their own rounding spread (L-curve ties: within a few search steps), and never past the
    # mixed-precision bar
    us = ~stable
    dist_cap = np.maximum(BAR_DIST, 4 * spread["dist"][us])
    mwf_cap = np.maximum(BAR_MWF, 4 * spread["mwf"][us])
    t = tie[us]
    dist_cap[t] = np.maximum(dist_cap[t], TIE_DIST)
    mwf_cap[t] = np.maximum(mwf_cap[t], 1e-3)
    assert np.all(rel[us] <= dist_cap), s
    assert np.all(mwf[us] <= np.minimum(mwf_cap, 1e-3)), s
    if mu_rel is not None:
        assert np.all(mu_rel[us & tie] <= TIE_MU), s
    if extra_check is not None:
        extra_check(stable)
    return s, stable
