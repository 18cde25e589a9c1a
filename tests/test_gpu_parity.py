"""GPU parity: the gfx950 path through the C ABI against the FP64 oracle on identical seeded inputs.

Every voxel is checked. The bar of the plain FP64 path (alpha <= 0.01 deg, distribution relative L2 <= 1e-6, MWF
(sfr) <= 1e-6, derived maps <= 1e-6 relative) holds on every voxel that the oracle itself computes stably at the
rounding level; voxels whose regularization search ends on a rounding-level tie are listed by id and bounded by the
oracle's own spread under ulp-scaled inputs (parity_util).
"""
import json
import os
from pathlib import Path

import numpy as np
import pytest

import parity_util as PU

pytestmark = pytest.mark.gpu

OUT = PU.OUT

CASES = {
    # name: (shape, nTE, TE, nT2, T2Range, Reg, nRefAngles, extra options, ncomp, mask fraction)
    "c1_chi2": ((16, 16, 2), 32, 10e-3, 40, (10e-3, 2.0), "chi2", 32, dict(Chi2Factor=1.02), 3, 1.0),
    "c2_gcv": ((12, 12, 3), 48, 10e-3, 40, (10e-3, 2.0), "gcv", 32, {}, 3, 0.6),
    "c3_lcurve": ((10, 10, 2), 56, 7.9e-3, 60, (8e-3, 2.0), "lcurve", 64, {}, 4, 0.6),
    "none": ((10, 10, 2), 32, 10e-3, 40, (10e-3, 2.0), "none", 32, {}, 3, 1.0),
    "mdp": ((10, 10, 2), 48, 10e-3, 40, (10e-3, 2.0), "mdp", 32, dict(NoiseLevel=2e-3), 3, 1.0),
    "m24_gcv_n40": ((8, 8, 2), 24, 7e-3, 40, (10e-3, 2.0), "gcv", 64, {}, 3, 1.0),
}


def _oracle_opts(oracle, shape, m, TE, n, T2R, reg, K, Kmin, extra, over):
    return oracle.make_t2map_opts(shape, m, TE, n, T2R, reg, nRefAngles=K, nRefAnglesMin=Kmin,
                                  Chi2Factor=extra.get("Chi2Factor"), NoiseLevel=extra.get("NoiseLevel"),
                                  RefConAngle=over.get("RefConAngle", 180.0), SetFlipAngle=over.get("SetFlipAngle"))


def _run(dmi, oracle, name, seed=11, **over):
    shape, m, TE, n, T2R, reg, K, extra, ncomp, frac = CASES[name]
    img, nin = dmi.synth_mse(shape, m, TE, ncomp=ncomp, mask_fraction=frac, seed=seed)
    kw = dict(TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K, SaveRegParam=True, SaveResidualNorm=True,
              SaveDecayCurve=True, Silent=True, **extra)
    kw.update(over)
    opts = dmi.T2mapOptions.from_image(img, **kw)
    maps, dist = dmi.T2mapSEcorr(img, opts)
    oo = _oracle_opts(oracle, shape, m, TE, n, T2R, reg, K, opts.nRefAnglesMin, extra, over)
    omaps, odist, nv = oracle.t2map(img, oo, nthreads=os.cpu_count())
    return img, opts, oo, maps, dist, omaps, odist, nv


def _curves_check(maps, omaps, sel):
    def f(stable):
        for k in ("resnorm", "chi2factor"):
            a, b = maps[k][sel][stable], omaps[k][sel][stable]
            err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            assert np.all(err <= PU.MAPS_REL), (k, float(err.max(initial=0.0)))
        a, b = maps["decaycurve"][sel][stable], omaps["decaycurve"][sel][stable]
        err = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
        assert np.all(err <= PU.MAPS_REL), ("decaycurve", float(err.max(initial=0.0)))
    return f


@pytest.mark.parametrize("name", list(CASES))
def test_t2map_matches_oracle(dmi, oracle, name):
    print(f"[parity] case={name}", flush=True)
    img, opts, oo, maps, dist, omaps, odist, nv = _run(dmi, oracle, name)
    T2R = CASES[name][4]
    # identical voxel selection and NaN placement
    assert np.array_equal(np.isnan(maps["gdn"]), np.isnan(omaps["gdn"]))
    assert np.array_equal(np.isnan(dist), np.isnan(odist))
    _, _, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R)
    s, stable = PU.check(name, oracle, T2R, sel, spread, maps, dist, omaps, odist,
                         extra_check=_curves_check(maps, omaps, sel))
    # the flip-angle search and the unregularized/chi2/mdp solves are rounding-stable on every voxel here
    if opts.Reg in ("none", "chi2", "mdp"):
        assert stable.all(), s


def test_decaybasisset_matches_oracle(dmi, oracle):
    opts = dmi.T2mapOptions(MatrixSize=(1, 1, 1), nTE=48, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="gcv",
                            nRefAngles=32, Silent=True)
    g = dmi.decaybasisset(opts)
    oo = oracle.make_t2map_opts((1, 1, 1), 48, 10e-3, 40, (10e-3, 2.0), "gcv", nRefAngles=32)
    img = np.zeros((1, 1, 1, 48), order="F")
    omaps, _, _ = oracle.t2map(img, oo, nthreads=1)
    assert g.shape == (48, 40, 32)
    assert np.max(np.abs(g - omaps["decaybasisset"])) <= 1e-13


def test_general_refcon_angle(dmi, oracle):
    print("[parity] case=refcon172", flush=True)
    img, opts, oo, maps, dist, omaps, odist, nv = _run(dmi, oracle, "none", RefConAngle=172.0)
    _, _, sel, spread = PU.oracle_ensemble(oracle, img, oo, (10e-3, 2.0))
    s, stable = PU.check("refcon172", oracle, (10e-3, 2.0), sel, spread, maps, dist, omaps, odist,
                         extra_check=_curves_check(maps, omaps, sel))
    assert stable.all(), s


def test_set_flip_angle(dmi, oracle):
    print("[parity] case=setflip", flush=True)
    img, opts, oo, maps, dist, omaps, odist, nv = _run(dmi, oracle, "c1_chi2", SetFlipAngle=165.0)
    _, _, sel, spread = PU.oracle_ensemble(oracle, img, oo, (10e-3, 2.0))
    assert np.all(maps["alpha"][sel] == 165.0)
    s, stable = PU.check("setflip165", oracle, (10e-3, 2.0), sel, spread, maps, dist, omaps, odist,
                         extra_check=_curves_check(maps, omaps, sel))
    assert stable.all(), s


def test_b1_map(dmi, oracle):
    print("[parity] case=b1map", flush=True)
    shape, m, TE, n, T2R = (8, 8, 2), 32, 10e-3, 40, (10e-3, 2.0)
    img, _ = dmi.synth_mse(shape, m, TE, mask_fraction=0.7, seed=5)
    b1 = np.asfortranarray(np.random.default_rng(0).uniform(120, 180, shape))
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg="none", nRefAngles=32, Silent=True,
                                       SaveRegParam=True)
    maps, dist = dmi.T2mapSEcorr(img, opts, alpha=b1)
    oo = oracle.make_t2map_opts(shape, m, TE, n, T2R, "none", nRefAngles=32)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R, alpha_in=b1)
    # alpha echoes the B1 map everywhere (also below threshold, as load_B1map! wrote it)
    assert np.array_equal(maps["alpha"], b1)
    s, stable = PU.check("b1map", oracle, T2R, sel, spread, maps, dist, omaps, odist)
    assert stable.all(), s


COMBOS = {
    # name: (base case, option overrides, B1 map)
    "setflip150_lcurve": ("c3_lcurve", dict(SetFlipAngle=150.0), False),
    "refcon160_gcv": ("c2_gcv", dict(RefConAngle=160.0), False),
    "minref90_t1_1p5_chi2": ("c1_chi2", dict(MinRefAngle=90.0, T1=1.5), False),
    "threshold_q40_chi2": ("c1_chi2", dict(Threshold="q40"), False),
    "k16_kmin3_mdp": ("mdp", dict(nRefAngles=16, nRefAnglesMin=3), False),
    "b1_lcurve": ("c3_lcurve", {}, True),
}


@pytest.mark.parametrize("name", list(COMBOS))
def test_option_combinations_match_oracle(dmi, oracle, name):
    """Options in combinations the configurations do not cover, every voxel against the oracle: a fixed flip angle
    with the L-curve, RefConAngle != 180 with GCV, another MinRefAngle and T1, a positive threshold, a short angle
    grid with a small nRefAnglesMin, and a B1 map with the L-curve. The oracle's options are taken field by field
    from the same T2mapOptions."""
    print(f"[parity] combo={name}", flush=True)
    base, over, b1 = COMBOS[name]
    shape, m, TE, n, T2R, reg, K, extra, ncomp, frac = CASES[base]
    img, _ = dmi.synth_mse(shape, m, TE, ncomp=ncomp, mask_fraction=frac, seed=29)
    over = dict(over)
    if over.get("Threshold") == "q40":
        e1 = img[..., 0]
        over["Threshold"] = float(np.quantile(e1[e1 > 0], 0.4))
    kw = dict(TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K, SaveRegParam=True, SaveResidualNorm=True,
              SaveDecayCurve=True, Silent=True, **extra)
    kw.update(over)
    opts = dmi.T2mapOptions.from_image(img, **kw)
    alpha = np.asfortranarray(np.random.default_rng(4).uniform(130, 180, shape)) if b1 else None
    maps, dist = dmi.T2mapSEcorr(img, opts, alpha=alpha)
    oo = oracle.make_t2map_opts(shape, m, TE, n, T2R, reg, T1=opts.T1, Threshold=opts.Threshold,
                                MinRefAngle=opts.MinRefAngle, nRefAngles=opts.nRefAngles,
                                nRefAnglesMin=opts.nRefAnglesMin, Chi2Factor=opts.Chi2Factor,
                                NoiseLevel=opts.NoiseLevel, RefConAngle=opts.RefConAngle,
                                SetFlipAngle=opts.SetFlipAngle)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R, alpha_in=alpha)
    assert np.array_equal(np.isnan(maps["gdn"]), ~sel) and sel.sum() > 0
    if "Threshold" in over:
        e1 = img[..., 0]
        assert np.array_equal(sel, e1 > over["Threshold"]) and 0 < sel.sum() < (e1 > 0).sum()
    if b1:
        assert np.array_equal(maps["alpha"], alpha)
    if opts.SetFlipAngle is not None:
        assert np.all(maps["alpha"][sel] == opts.SetFlipAngle)
    s, stable = PU.check(name, oracle, T2R, sel, spread, maps, dist, omaps, odist,
                         extra_check=_curves_check(maps, omaps, sel))
    if reg in ("chi2", "mdp", "none"):
        assert stable.all(), s


def test_t2part_matches_oracle(dmi, oracle):
    rng = np.random.default_rng(3)
    dist = np.asfortranarray(rng.uniform(0, 1, (9, 7, 3, 40)) * (rng.uniform(0, 1, (9, 7, 3, 40)) > 0.5))
    dist[0, 0, 0, 5] = np.nan
    dist[1, 0, 0, :] = 0.0
    for sig in (None, 5e-3):
        g = dmi.T2partSEcorr(dist, T2Range=(10e-3, 2.0), SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0), Sigmoid=sig,
                             Silent=True)
        o = oracle.t2part(dist, (10e-3, 2.0), (10e-3, 40e-3), (40e-3, 2.0), Sigmoid=sig)
        for k in ("sfr", "sgm", "mfr", "mgm"):
            assert np.array_equal(np.isnan(g[k]), np.isnan(o[k])), k
            ok = ~np.isnan(o[k])
            assert np.max(np.abs(g[k][ok] - o[k][ok]) / np.maximum(np.abs(o[k][ok]), 1e-300)) <= 1e-12, k


def test_threshold_edge_cases(dmi):
    shape = (4, 4, 1)
    img = np.zeros((*shape, 32), order="F")
    opts = dmi.T2mapOptions.from_image(img, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="none", Silent=True)
    maps, dist = dmi.T2mapSEcorr(img, opts)
    assert np.all(np.isnan(maps["gdn"])) and np.all(np.isnan(dist))
    img2, _ = dmi.synth_mse(shape, 32, 10e-3, mask_fraction=1.0, seed=1)
    img2[0, 0, 0, 3] = np.nan  # a NaN echo propagates to that voxel's outputs, others unaffected
    maps2, dist2 = dmi.T2mapSEcorr(img2, opts)
    assert np.sum(~np.isnan(maps2["gdn"])) >= shape[0] * shape[1] - 1


@pytest.mark.parametrize("name", ["c3_lcurve", "c2_gcv"])
def test_mixed_precision_error(dmi, oracle, name):
    """FP32 EPG (grid bases and the voxel basis) with FP64 NNLS and regularization, against the FP64 oracle:
    alpha <= 0.1 deg and MWF <= 1e-3 absolute per voxel; every output's error distribution is written out."""
    print(f"[parity] case=mixed_{name}", flush=True)
    shape, m, TE, n, T2R, reg, K, extra, ncomp, frac = CASES[name]
    img, opts, oo, maps, dist, omaps, odist, nv = _run(dmi, oracle, name, precision="mixed")
    mask = ~np.isnan(omaps["gdn"])
    rel = np.linalg.norm(dist[mask] - odist[mask], axis=1) / np.linalg.norm(odist[mask], axis=1)
    da = np.abs(maps["alpha"][mask] - omaps["alpha"][mask])
    pg = oracle.t2part(dist, T2R, (T2R[0], 40e-3), (40e-3, T2R[1]))
    po = oracle.t2part(odist, T2R, (T2R[0], 40e-3), (40e-3, T2R[1]))
    mwf = np.abs(pg["sfr"][mask] - po["sfr"][mask])
    sgm = np.abs(pg["sgm"][mask] - po["sgm"][mask]) / np.abs(po["sgm"][mask])
    chi = np.abs(maps["chi2factor"][mask] - omaps["chi2factor"][mask])

    def q(x):
        return dict(max=float(np.nanmax(x)), p99=float(np.nanquantile(x, 0.99)), median=float(np.nanmedian(x)))

    s = dict(case=f"mixed_{name}", voxels=int(mask.sum()), alpha_abs_deg=q(da), dist_rel_l2=q(rel), mwf_abs=q(mwf),
             sgm_rel=q(sgm), chi2factor_abs=q(chi))
    if OUT:
        Path(OUT).mkdir(parents=True, exist_ok=True)
        with open(Path(OUT) / f"parity_mixed_{name}.json", "w") as f:
            json.dump(s, f, indent=1)
    print(json.dumps(s), flush=True)
    assert np.max(da) <= 0.1, s
    assert np.max(mwf) <= 1e-3, s


@pytest.mark.parametrize("shape,m,n,K,reg", [
    ((3, 2, 1), 124, 64, 64, "gcv"),       # device maximums: 124 echoes, 64 T2 bins, 64 angles (3 rows per lane)
    ((2, 2, 1), 124, 64, 64, "lcurve"),
    ((4, 3, 1), 4, 2, 2, "chi2"),          # reference minimums: 4 echoes, 2 T2 bins, 2 angles
    ((1, 1, 1), 32, 40, 32, "none"),       # a single voxel
    ((5, 1, 3), 20, 64, 16, "mdp"),        # n > m
])
def test_size_limits(dmi, oracle, shape, m, n, K, reg):
    print(f"[parity] case=limits {shape} m={m} n={n} K={K} {reg}", flush=True)
    TE, T2R = 10e-3, (10e-3, 2.0)
    img, _ = dmi.synth_mse(shape, m, TE, mask_fraction=1.0, seed=17)
    extra = dict(Chi2Factor=1.02) if reg == "chi2" else dict(NoiseLevel=1e-3) if reg == "mdp" else {}
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K,
                                       nRefAnglesMin=min(5, K), Silent=True, SaveRegParam=True, **extra)
    maps, dist = dmi.T2mapSEcorr(img, opts)
    oo = oracle.make_t2map_opts(shape, m, TE, n, T2R, reg, nRefAngles=K, nRefAnglesMin=min(5, K),
                                Chi2Factor=extra.get("Chi2Factor"), NoiseLevel=extra.get("NoiseLevel"))
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R)
    assert np.array_equal(np.isnan(dist), np.isnan(odist))
    s, stable = PU.check(f"limits_m{m}_n{n}_K{K}_{reg}", oracle, T2R, sel, spread, maps, dist, omaps, odist)
    if reg in ("none", "chi2", "mdp"):
        assert stable.all(), s


@pytest.mark.parametrize("m,reg,setflip", [(33, "gcv", None), (25, "chi2", 165.0)])
def test_odd_echo_count_basis_layout(dmi, oracle, m, reg, setflip):
    """An odd echo count pads every column of the device bases to an even stride (grid bases, the fixed basis and
    the voxel's basis). Distributions and angles match the oracle; the saved per-voxel basis equals the oracle's
    fixed basis (SetFlipAngle) and reproduces the saved decay curve from the distribution (both paths)."""
    print(f"[parity] case=odd_m m={m} {reg} setflip={setflip}", flush=True)
    shape, TE, n, K, T2R = (4, 3, 2), 10e-3, 40, 32, (10e-3, 2.0)
    img, _ = dmi.synth_mse(shape, m, TE, mask_fraction=1.0, seed=23)
    extra = dict(Chi2Factor=1.02) if reg == "chi2" else {}
    if setflip is not None:
        extra["SetFlipAngle"] = setflip
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg=reg, nRefAngles=K, nRefAnglesMin=5,
                                       Silent=True, SaveRegParam=True, SaveDecayCurve=True, SaveNNLSBasis=True,
                                       **extra)
    maps, dist = dmi.T2mapSEcorr(img, opts)
    oo = oracle.make_t2map_opts(shape, m, TE, n, T2R, reg, nRefAngles=K, nRefAnglesMin=5,
                                Chi2Factor=extra.get("Chi2Factor"), SetFlipAngle=setflip)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R)
    assert np.array_equal(np.isnan(dist), np.isnan(odist))
    mask = sel
    assert mask.sum() == np.prod(shape)
    PU.check(f"odd_m{m}_{reg}", oracle, T2R, sel, spread, maps, dist, omaps, odist)
    if setflip is not None:
        # one shared basis (m, n), copied out of the padded device layout
        assert maps["decaybasis"].shape == (m, n)
        assert np.max(np.abs(maps["decaybasis"] - omaps["decaybasisset"])) <= 1e-13
        basis = np.broadcast_to(maps["decaybasis"], (int(mask.sum()), m, n))
    else:
        basis = maps["decaybasis"][mask]                  # (voxels, m, n)
        assert basis.shape[1:] == (m, n)
    fit = np.einsum("vin,vn->vi", basis, dist[mask])
    curve = maps["decaycurve"][mask]
    assert np.max(np.abs(fit - curve)) <= 1e-12 * np.max(np.abs(curve))


def test_float32_inputs(dmi):
    """Float32 arrays: outputs are the FP64 path's on the widened input, rounded once to Float32."""
    print("[parity] case=float32", flush=True)
    img, _ = dmi.synth_mse((6, 5, 2), 32, 10e-3, mask_fraction=0.8, seed=21)
    img32 = img.astype(np.float32, order="F")
    opts = dmi.T2mapOptions.from_image(img32, TE=10e-3, nT2=40, T2Range=(10e-3, 2.0), Reg="chi2", Chi2Factor=1.02,
                                       nRefAngles=32, Silent=True, SaveRegParam=True)
    maps32, dist32 = dmi.T2mapSEcorr(img32, opts)
    maps64, dist64 = dmi.T2mapSEcorr(img32.astype(np.float64, order="F"), opts)
    assert dist32.dtype == np.float32 and maps32["gdn"].dtype == np.float32
    np.testing.assert_array_equal(dist32, dist64.astype(np.float32))
    for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha", "mu", "chi2factor"):
        np.testing.assert_array_equal(maps32[k], maps64[k].astype(np.float32))
    popts = dmi.T2partOptions.from_t2map(opts, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0))
    p32 = dmi.T2partSEcorr(dist32, popts)
    p64 = dmi.T2partSEcorr(dist32.astype(np.float64, order="F"), popts)
    for k in ("sfr", "sgm", "mfr", "mgm"):
        assert p32[k].dtype == np.float32
        np.testing.assert_array_equal(p32[k], p64[k].astype(np.float32))


def test_full_volume_sampled(dmi, oracle):
    """The 240x240x48 gcv volume (1.66M voxels above threshold) through the host-array entry point: every output
    against the oracle on a seeded sample of in-mask voxels spread over the whole volume (so over the whole work
    queue of the persistent grid), plus size-independent invariants on all voxels."""
    shape, m, TE, n, T2R = (240, 240, 48), 48, 10e-3, 40, (10e-3, 2.0)
    print(f"[parity] full volume {shape} nTE={m} gcv", flush=True)
    img, nin = dmi.synth_mse(shape, m, TE, ncomp=3, mask_fraction=0.6, seed=3)
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=T2R, Reg="gcv", nRefAngles=32,
                                       SaveRegParam=True, Silent=True)
    maps, dist = dmi.T2mapSEcorr(img, opts)
    inmask = img[..., 0] > 0
    assert int(inmask.sum()) == nin
    # NaN exactly outside the threshold mask; finite, positive totals inside; gdn is the distribution's sum
    assert np.array_equal(~np.isnan(maps["gdn"]), inmask)
    assert np.array_equal(~np.isnan(dist[..., 0]), inmask)
    g = maps["gdn"][inmask]
    assert np.all(np.isfinite(g)) and np.all(g > 0)
    d = dist[inmask]
    assert np.all(d >= 0)
    assert np.max(np.abs(d.sum(axis=1) - g) / g) <= 1e-12
    a = maps["alpha"][inmask]
    assert np.all((a >= 50.0 - 1e-9) & (a <= 180.0 + 1e-9))
    parts = dmi.T2partSEcorr(dist, dmi.T2partOptions.from_dist(dist, T2Range=T2R, SPWin=(T2R[0], 40e-3),
                                                              MPWin=(40e-3, T2R[1]), Silent=True))
    sfr = parts["sfr"][inmask]
    assert np.all((sfr >= 0) & (sfr <= 1))
    # sampled parity against the oracle
    rng = np.random.default_rng(1234)
    lin = np.flatnonzero(inmask.ravel(order="F"))
    vox = np.sort(rng.choice(lin, size=3000, replace=False))
    oo = oracle.make_t2map_opts(shape, m, TE, n, T2R, "gcv", nRefAngles=32, nRefAnglesMin=opts.nRefAnglesMin)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R, vox=vox)
    s, stable = PU.check("c2_full_volume_sampled", oracle, T2R, sel, spread, maps, dist, omaps, odist)
    # rounding-level ties of the GCV search are rare
    assert (~stable).sum() <= 0.02 * stable.size, s


def test_c3_volume_plain_and_mixed(dmi, oracle):
    """The 240x240x113, 56-echo, nT2 60, L-curve, 64-angle volume of the headline (3.9M voxels above threshold),
    plain FP64 and mixed precision, through the host-array entry point: invariants on every voxel, the plain path
    against the oracle on a seeded sample of 1500 in-mask voxels (per-voxel bar, rounding ties bounded), and the
    mixed path against the plain one on every voxel: MWF within 1e-3 on at least 99.99% of voxels and never more
    than 0.05, alpha within 0.1 degree."""
    shape, m, TE, n, T2R = (240, 240, 113), 56, 7.9e-3, 60, (8e-3, 2.0)
    print(f"[parity] C3 volume {shape} nTE={m} lcurve", flush=True)
    img, nin = dmi.synth_mse(shape, m, TE, ncomp=4, mask_fraction=0.6, seed=2024)
    kw = dict(TE=TE, nT2=n, T2Range=T2R, Reg="lcurve", nRefAngles=64, SaveRegParam=True, Silent=True)
    maps, dist = dmi.T2mapSEcorr(img, dmi.T2mapOptions.from_image(img, **kw))
    inmask = img[..., 0] > 0
    assert int(inmask.sum()) == nin
    assert np.array_equal(~np.isnan(maps["gdn"]), inmask)
    g = maps["gdn"][inmask]
    assert np.all(np.isfinite(g)) and np.all(g > 0)
    assert np.max(np.abs(dist[inmask].sum(axis=1) - g) / g) <= 1e-12
    # the sampled voxels packed into a (N, 1, 1) image for the oracle (voxels are independent of their position)
    rng = np.random.default_rng(4321)
    lin = np.flatnonzero(inmask.ravel(order="F"))
    vox = np.sort(rng.choice(lin, size=1500, replace=False))
    idx = np.unravel_index(vox, shape, order="F")
    N = len(vox)
    pimg = np.asfortranarray(img[idx][:, None, None, :])
    pmaps = {k: np.asfortranarray(v[idx][:, None, None]) for k, v in maps.items() if v.shape == shape}
    pdist = np.asfortranarray(dist[idx][:, None, None, :])
    oo = oracle.make_t2map_opts((N, 1, 1), m, TE, n, T2R, "lcurve", nRefAngles=64)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, pimg, oo, T2R)
    assert sel.all()
    s, stable = PU.check("c3_volume_sampled", oracle, T2R, sel, spread, pmaps, pdist, omaps, odist)
    spw = dict(SPWin=(T2R[0], 40e-3), MPWin=(40e-3, T2R[1]), Silent=True)
    sfr = dmi.T2partSEcorr(dist, dmi.T2partOptions.from_dist(dist, T2Range=T2R, **spw))["sfr"][inmask]
    alpha = maps["alpha"][inmask]
    del maps, dist
    mmaps, mdist = dmi.T2mapSEcorr(img, dmi.T2mapOptions.from_image(img, precision="mixed", **kw))
    msfr = dmi.T2partSEcorr(mdist, dmi.T2partOptions.from_dist(mdist, T2Range=T2R, **spw))["sfr"][inmask]
    err = np.abs(msfr - sfr)
    da = np.abs(mmaps["alpha"][inmask] - alpha)
    big = da > 0.1
    print(f"[parity] mixed vs plain on {err.size} voxels: MWF max {err.max():.3e}, > 1e-3: {(err > 1e-3).sum()}; "
          f"alpha max {da.max():.3e} deg, > 0.1 deg: {big.sum()}, > 1e-3 deg: {(da > 1e-3).sum()}, "
          f"MWF of those > 0.1 deg: max {err[big].max(initial=0):.3e}, p50 {np.median(err[big]) if big.any() else 0:.3e}; "
          f"alpha p99.99 {np.quantile(da, 0.9999):.3e}", flush=True)
    assert (err > 1e-3).mean() <= 1e-4 and err.max() <= 0.05
    assert (da > 0.1).mean() <= 1e-4


@pytest.mark.parametrize("name,shape,m,n,K,reg,extra", [
    ("c3_lcurve", (10, 10, 2), 56, 60, 64, "lcurve", {}),
    ("c1_chi2", (12, 12, 2), 32, 40, 32, "chi2", dict(Chi2Factor=1.02)),
    ("mdp", (10, 10, 2), 48, 40, 32, "mdp", dict(NoiseLevel=2e-3)),
    ("chi2_m56_n60", (8, 8, 2), 56, 60, 64, "chi2", dict(Chi2Factor=1.02)),
    ("mdp_m64_n60", (6, 6, 2), 64, 60, 32, "mdp", dict(NoiseLevel=2e-3)),
    ("limits_lcurve", (3, 2, 1), 124, 64, 64, "lcurve", {}),
    ("c2_gcv", (8, 8, 2), 48, 40, 32, "gcv", {}),  # no split schedule for GCV: "split" runs the all-LDS one
])
def test_split_schedule_is_bit_identical(dmi, name, shape, m, n, K, reg, extra):
    """The split schedule (Tikhonov rows beyond an LDS cutoff kept in the workgroup's global slot, so that more
    workgroups share a CU) performs the same arithmetic in the same order as the all-LDS schedule: every output is
    bit-identical, and the split launch really holds fewer rows in LDS with more workgroups per CU."""
    print(f"[schedule] case={name}", flush=True)
    T2R = (8e-3, 2.0) if m == 56 else (10e-3, 2.0)
    img, _ = dmi.synth_mse(shape, m, 7.9e-3 if m == 56 else 10e-3, ncomp=4 if m == 56 else 3, mask_fraction=1.0,
                           seed=29)
    res = {}
    for sched in ("lds", "split"):
        opts = dmi.T2mapOptions.from_image(img, TE=7.9e-3 if m == 56 else 10e-3, nT2=n, T2Range=T2R, Reg=reg,
                                           nRefAngles=K, SaveRegParam=True, SaveResidualNorm=True, Silent=True,
                                           schedule=sched, **extra)
        maps, dist = dmi.T2mapSEcorr(img, opts)
        st = dmi.api._lib.context(None).stats()
        res[sched] = (maps, dist, st)
    (ml, dl, sl), (ms, ds, ss) = res["lds"], res["split"]
    print(f"[schedule] lds rows {sl['lds_rows']} x {sl['workgroups_per_cu']}/CU, "
          f"split rows {ss['lds_rows']} x {ss['workgroups_per_cu']}/CU", flush=True)
    assert np.array_equal(dl, ds, equal_nan=True)
    for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha", "mu", "chi2factor", "resnorm"):
        assert np.array_equal(ml[k], ms[k], equal_nan=True), k
    assert sl["lds_rows"] == m + n
    if reg == "gcv" or sl["workgroups_per_cu"] >= 4:
        # no split for GCV; none either where the all-LDS schedule already has one workgroup per SIMD
        assert ss["lds_rows"] == m + n and ss["workgroups_per_cu"] == sl["workgroups_per_cu"]
    else:
        assert ss["lds_rows"] < m + n and ss["workgroups_per_cu"] > sl["workgroups_per_cu"]
