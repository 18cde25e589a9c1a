"""GPU: the command line layer end to end (as the reference's test/cli.jl): requested files only, settings files
copied, results equal to the API's on the same input, T2part-only runs on a distribution, masks and B1 maps."""
import numpy as np
import pytest
import scipy.io

pytestmark = pytest.mark.gpu

ARGS = ["--TE", "0.01", "--nT2", "40", "--T2Range", "0.01", "2.0", "--SPWin", "0.01", "0.04", "--MPWin", "0.04", "2.0",
        "--nRefAngles", "32", "--SaveRegParam"]


def _image(dmi, seed=3):
    img, _ = dmi.synth_mse((6, 5, 2), 32, 10e-3, mask_fraction=0.8, seed=seed)
    return img


@pytest.mark.parametrize("fmt,settings", [(".mat", False), ("v7.3 .mat", False), (".nii.gz", True), (".par", False)])
def test_cli_matches_api(dmi, tmp_path, fmt, settings):
    from decaes_amd import cli, fileio, mat73
    print(f"[cli] {fmt} settings={settings}", flush=True)
    img = _image(dmi)
    inp = tmp_path / ("scan" + fmt.split()[-1])
    if fmt == ".mat":
        scipy.io.savemat(inp, {"img": img})
    elif fmt == "v7.3 .mat":
        mat73.write(inp, {"img": img})
    elif fmt == ".par":
        # 16-bit pixel values with a power-of-two scale slope, so the floating-point values are exact
        import parrec_files as PF
        step = 2.0 ** np.ceil(np.log2(img.max() / 60000.0))
        pv = np.round(img / step).astype(np.uint16)
        img = pv.astype(np.float64) * step
        PF.write_rec(tmp_path / "scan.rec", PF.write_par(inp, pv[..., None], 1.0, 0.0, 1.0 / step))
    else:
        fileio.write_nifti(inp, img)
    out = tmp_path / "out"
    args = [str(inp), "--T2map", "--T2part", "-o", str(out), "-q", "--Reg", "chi2", "--RegParams", "1.02"] + ARGS
    if settings:
        s = tmp_path / "settings.txt"
        s.write_text("\n".join(args))
        args = ["@" + str(s)]
    assert cli.main(args) == 0
    for ext in (".t2dist.mat", ".t2maps.mat", ".t2parts.mat", ".log"):
        assert (out / ("scan" + ext)).is_file(), ext
    for ext in (".t2dist.mat", ".t2maps.mat", ".t2parts.mat"):
        assert mat73.is_mat73(out / ("scan" + ext)), ext  # MAT v7.3, as MAT.matwrite writes
    assert (out / "scan.settings.txt").is_file() == settings
    opts = dmi.T2mapOptions.from_image(img, TE=0.01, nT2=40, T2Range=(0.01, 2.0), Reg="chi2", Chi2Factor=1.02,
                                       nRefAngles=32, SaveRegParam=True, Silent=True)
    maps, dist = dmi.T2mapSEcorr(img, opts)
    parts = dmi.T2partSEcorr(dist, dmi.T2partOptions.from_t2map(opts, SPWin=(0.01, 0.04), MPWin=(0.04, 2.0)))
    np.testing.assert_array_equal(fileio.read_mat(out / "scan.t2dist.mat")["dist"], dist)
    m = fileio.read_mat(out / "scan.t2maps.mat")
    assert set(k for k in m if not k.startswith("__")) == set(maps)
    for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha", "mu", "chi2factor", "decaybasisset"):
        np.testing.assert_array_equal(m[k], maps[k])
    p = fileio.read_mat(out / "scan.t2parts.mat")
    for k in ("sfr", "sgm", "mfr", "mgm"):
        np.testing.assert_array_equal(p[k], parts[k])


def test_cli_t2part_only_mask_b1(dmi, tmp_path):
    from decaes_amd import cli, fileio
    print("[cli] t2part-only, mask, B1 map", flush=True)
    img = _image(dmi, seed=4)
    mask = np.ones(img.shape[:3])
    mask[0, :, :] = 0.0
    b1 = np.full(img.shape[:3], 165.0)
    scipy.io.savemat(tmp_path / "scan.mat", {"img": img})
    fileio.write_nifti(tmp_path / "mask.nii", mask)
    fileio.write_nifti(tmp_path / "b1.nii", b1)
    out = tmp_path / "o"
    assert cli.main([str(tmp_path / "scan.mat"), "--T2map", "-o", str(out), "-q", "--Reg", "none", "-m",
                     str(tmp_path / "mask.nii"), "--B1map", str(tmp_path / "b1.nii")] + ARGS) == 0
    assert not (out / "scan.t2parts.mat").exists()
    m = fileio.read_mat(out / "scan.t2maps.mat")
    assert np.all(np.isnan(m["gdn"][0])) and np.all(m["alpha"] == 165.0)
    opts = dmi.T2mapOptions.from_image(img * mask[..., None], TE=0.01, nT2=40, T2Range=(0.01, 2.0), Reg="none",
                                       nRefAngles=32, SaveRegParam=True, Silent=True)
    maps, dist = dmi.T2mapSEcorr(img * mask[..., None], opts, alpha=b1)
    np.testing.assert_array_equal(fileio.read_mat(out / "scan.t2dist.mat")["dist"], dist)
    # T2part only: the input is the distribution; nT2 is inferred from it
    scipy.io.savemat(tmp_path / "d.mat", {"dist": dist})
    assert cli.main([str(tmp_path / "d.mat"), "--T2part", "-o", str(out), "-q", "--T2Range", "0.01", "2.0",
                     "--SPWin", "0.01", "0.04", "--MPWin", "0.04", "2.0", "--Sigmoid", "0.005"]) == 0
    p = fileio.read_mat(out / "d.t2parts.mat")
    ref = dmi.T2partSEcorr(dist, T2Range=(0.01, 2.0), SPWin=(0.01, 0.04), MPWin=(0.04, 2.0), Sigmoid=0.005)
    for k in ("sfr", "sgm", "mfr", "mgm"):
        np.testing.assert_array_equal(p[k], ref[k])
