"""The per-voxel parity checker (parity_util) on the CPU: the oracle passes against itself, a planted error on one
voxel fails, and the rounding ties it exempts are real (the oracle's own result moves by ~1e-4 on them when its
input changes in the last bit)."""
import numpy as np
import pytest

import parity_util as PU


def _case(dmi, oracle, reg="gcv", shape=(6, 6, 2), m=48, n=40, K=32, seed=11):
    T2R = (10e-3, 2.0)
    img, _ = dmi.synth_mse(shape, m, 10e-3, mask_fraction=1.0, seed=seed)
    oo = oracle.make_t2map_opts(shape, m, 10e-3, n, T2R, reg, nRefAngles=K,
                                Chi2Factor=1.02 if reg == "chi2" else None)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R, nthreads=8)
    return T2R, omaps, odist, sel, spread


def test_oracle_passes_against_itself(dmi, oracle):
    T2R, omaps, odist, sel, spread = _case(dmi, oracle, "chi2")
    s, stable = PU.check("cpu_self_chi2", oracle, T2R, sel, spread, dict(omaps), odist.copy(), omaps, odist)
    assert stable.all()


def test_planted_error_on_a_stable_voxel_fails(dmi, oracle):
    T2R, omaps, odist, sel, spread = _case(dmi, oracle, "chi2")
    bad = odist.copy()
    i = np.argwhere(sel)[3]
    bad[tuple(i)] *= 1 + 5e-6          # one voxel's distribution off by 5e-6 relative
    with pytest.raises(AssertionError):
        PU.check("cpu_planted", oracle, T2R, sel, spread, dict(omaps), bad, omaps, odist)
    badm = {k: v.copy() for k, v in omaps.items()}
    badm["ggm"][tuple(i)] *= 1 + 1e-5  # or one derived map
    with pytest.raises(AssertionError):
        PU.check("cpu_planted_ggm", oracle, T2R, sel, spread, badm, odist.copy(), omaps, odist)


def test_gcv_rounding_ties_are_the_oracles_own(dmi, oracle):
    """Three voxels of the 240x240x48 gcv volume whose GCV Brent search ends on a rounding-level tie: the oracle
    itself moves them by > 1e-5 when the image is scaled by 1 + 2^-52, and the ensemble marks them unstable; the
    singular values from Gram eigenvalues or from a bidiagonalization move them no further than that."""
    shape, m, TE, n, T2R = (240, 240, 48), 48, 10e-3, 40, (10e-3, 2.0)
    img, _ = dmi.synth_mse(shape, m, TE, ncomp=3, mask_fraction=0.6, seed=3)
    ties = np.array([1417228, 2272826, 1661388])
    calm = np.array([1245840, 1284481, 1632961])
    vox = np.sort(np.concatenate([ties, calm]))
    oo = oracle.make_t2map_opts(shape, m, TE, n, T2R, "gcv", nRefAngles=32, nRefAnglesMin=5)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, T2R, vox=vox, nthreads=8)
    lin = np.ravel_multi_index(np.nonzero(sel), sel.shape, order="F")
    assert np.array_equal(np.sort(lin), vox)
    tie = np.isin(lin, ties)
    assert np.all(spread["dist"][tie] > 1e-5) and np.all(spread["dist"][~tie] < 1e-9)
    for route in (1, 2):
        oracle.lib().orc_set_svd_route(route)
        try:
            m2, d2 = oracle.t2map_subset(img, oo, vox, nthreads=8)
        finally:
            oracle.lib().orc_set_svd_route(0)
        rel = np.linalg.norm(d2[sel] - odist[sel], axis=1) / np.linalg.norm(odist[sel], axis=1)
        assert np.all(rel[~tie] < 1e-7), route
        assert np.all(rel[tie] <= 4 * spread["dist"][tie]), route
    del img


def test_lcurve_final_choice_ties_are_marked(dmi, oracle):
    """The L-curve's final argmax margin (relative gap of the best and second-best cached curvature) from the oracle:
    a voxel of a 24-echo, 40-bin volume whose corner is a 8e-7 tie (the top two curvatures 15.7793421 and 15.7793289)
    is marked a tie; the margin is finite and positive on every voxel and most voxels are no tie."""
    shape, m, TE = (20, 20, 1), 24, 7e-3
    img, _ = dmi.synth_mse(shape, m, TE, ncomp=3, mask_fraction=0.5752288710445501, seed=642)
    oo = oracle.make_t2map_opts(shape, m, TE, 40, (10e-3, 2.0), "lcurve", nRefAngles=64)
    omaps, odist, sel, spread = PU.oracle_ensemble(oracle, img, oo, (10e-3, 2.0), nthreads=8)
    lin = np.ravel_multi_index(np.nonzero(sel), sel.shape, order="F")
    assert spread["tie"][lin == 257].all()
    assert 0 < spread["tie"].mean() < 0.6
