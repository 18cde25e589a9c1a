"""The C ABI driven from a compiled C host program (decaes_amd/dmi_run, built from csrc/dmi_run.c against
libdecaes_amd.so; no Python or torch in that process): it synthesizes a seeded image, calls dmi_t2map_f64 and
dmi_t2part_f64 on host arrays, and writes raw column-major float64 files. The GPU tests hold its outputs to the oracle
with the per-voxel bar, and to the Python API's outputs bit for bit (one library, one arithmetic)."""
import json
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import parity_util as PU

RUN = Path(__file__).resolve().parents[1] / "decaes.jl_amd" / "decaes_amd" / "dmi_run"

CASES = {
    # shape, nTE, TE, nT2, Reg, nRefAngles, seed, mask fraction
    "c_host_chi2": ((12, 10, 3), 32, 10e-3, 40, "chi2", 32, 11, 0.8),
    "c_host_lcurve": ((8, 8, 2), 48, 8e-3, 50, "lcurve", 24, 12, 0.9),
}


def _read(d, name, shape):
    return np.fromfile(d / f"{name}.f64", dtype="<f8").reshape(shape, order="F")


def test_c_host_usage():
    if not RUN.exists():
        pytest.skip("dmi_run not built")
    r = subprocess.run([str(RUN)], capture_output=True, text=True)
    assert r.returncode == 2 and "usage: dmi_run" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_c_host_matches_oracle_and_api(dmi, oracle, tmp_path, name):
    assert RUN.exists(), "dmi_run is built with the library"
    shape, m, TE, n, reg, K, seed, frac = CASES[name]
    print(f"[c-host] case={name}", flush=True)
    r = subprocess.run([str(RUN), str(tmp_path), *map(str, shape), str(m), repr(TE), str(n), reg, str(K), str(seed),
                        repr(frac)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    img = _read(tmp_path, "image", shape + (m,))
    ref_img, nin = dmi.synth_mse(shape, m, TE, ncomp=3, mask_fraction=frac, seed=seed)
    assert np.array_equal(img, ref_img) and info["voxels_in_mask"] == nin
    dist = _read(tmp_path, "dist", shape + (n,))
    maps = {k: _read(tmp_path, k, shape) for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha", "resnorm", "mu",
                                                   "chi2factor")}
    parts = {k: _read(tmp_path, k, shape) for k in ("sfr", "sgm", "mfr", "mgm")}

    # the same calls through the Python API: the same bits
    extra = dict(Chi2Factor=1.02) if reg == "chi2" else {}
    opts = dmi.T2mapOptions.from_image(img, TE=TE, nT2=n, T2Range=(10e-3, 2.0), Reg=reg, nRefAngles=K,
                                       SaveRegParam=True, SaveResidualNorm=True, Silent=True, **extra)
    amaps, adist = dmi.T2mapSEcorr(img, opts)
    assert np.array_equal(dist, adist, equal_nan=True)
    for k, v in maps.items():
        assert np.array_equal(v, amaps[k], equal_nan=True), k
    aparts = dmi.T2partSEcorr(adist, dmi.T2partOptions.from_t2map(opts, SPWin=(10e-3, 40e-3), MPWin=(40e-3, 2.0)))
    for k, v in parts.items():
        assert np.array_equal(v, aparts[k], equal_nan=True), k

    # against the oracle, every voxel
    oo = oracle.make_t2map_opts(shape, m, TE, n, (10e-3, 2.0), reg, nRefAngles=K, nRefAnglesMin=min(5, K),
                                Chi2Factor=extra.get("Chi2Factor"))
    omaps, odist, nv = oracle.t2map(img, oo, nthreads=os.cpu_count())
    sel = ~np.isnan(omaps["gdn"])
    assert np.array_equal(sel, ~np.isnan(maps["gdn"])) and nv == info["voxels"] == sel.sum() > 0
    _, _, sel2, spread = PU.oracle_ensemble(oracle, img, oo, (10e-3, 2.0))
    assert np.array_equal(sel, sel2)
    s, stable = PU.check(name, oracle, (10e-3, 2.0), sel, spread, maps, dist, omaps, odist)
    if reg == "chi2":
        assert stable.all(), s
