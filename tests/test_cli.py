"""CPU: the command line layer's file formats and argument handling (src/main.jl; test/cli.jl's conventions)."""
import struct
from pathlib import Path

import numpy as np
import pytest
import scipy.io

from decaes_amd import cli, fileio


@pytest.mark.parametrize("suffix", [".nii", ".nii.gz"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int16, np.uint8])
def test_nifti_roundtrip(tmp_path, suffix, dtype):
    rng = np.random.default_rng(0)
    a = np.asfortranarray((rng.random((5, 4, 3, 6)) * 100).astype(dtype))
    f = tmp_path / ("img" + suffix)
    fileio.write_nifti(f, a)
    b = fileio.read_nifti(f)
    assert b.shape == a.shape and b.dtype == np.float64
    np.testing.assert_array_equal(b, a.astype(np.float64))


def test_nifti_scaling_and_slope_zero(tmp_path):
    a = np.asfortranarray(np.arange(24, dtype=np.int16).reshape(2, 3, 4))
    f = tmp_path / "s.nii"
    fileio.write_nifti(f, a)
    raw = bytearray(f.read_bytes())
    struct.pack_into("<2f", raw, 112, 2.0, 1.0)
    f.write_bytes(bytes(raw))
    np.testing.assert_array_equal(fileio.read_nifti(f), 2.0 * a + 1.0)
    struct.pack_into("<2f", raw, 112, 0.0, 5.0)  # slope 0: unscaled (the reference's rule)
    f.write_bytes(bytes(raw))
    np.testing.assert_array_equal(fileio.read_nifti(f), a.astype(np.float64))


def test_nifti_big_endian(tmp_path):
    a = np.asfortranarray(np.arange(12, dtype=np.float32).reshape(3, 4))
    hdr = bytearray(352)
    struct.pack_into(">i", hdr, 0, 348)
    struct.pack_into(">8h", hdr, 40, 2, 3, 4, 1, 1, 1, 1, 1)
    struct.pack_into(">2h", hdr, 70, 16, 32)
    struct.pack_into(">f", hdr, 108, 352.0)
    (tmp_path / "be.nii").write_bytes(bytes(hdr) + a.astype(">f4").tobytes(order="F"))
    np.testing.assert_array_equal(fileio.read_nifti(tmp_path / "be.nii"), a.astype(np.float64))


def test_load_image_mat_and_ndims(tmp_path):
    img4 = np.asfortranarray(np.random.default_rng(1).random((4, 3, 2, 5)))
    scipy.io.savemat(tmp_path / "a.mat", {"zimg": img4, "mask3": np.ones((4, 3, 2))})
    np.testing.assert_array_equal(fileio.load_image(tmp_path / "a.mat", 4), img4)
    np.testing.assert_array_equal(fileio.load_image(tmp_path / "a.mat", 3), np.ones((4, 3, 2)))
    scipy.io.savemat(tmp_path / "two.mat", {"b": img4, "a": img4 * 2})
    with pytest.warns(UserWarning, match="Multiple possible images"):
        np.testing.assert_array_equal(fileio.load_image(tmp_path / "two.mat", 4), img4 * 2)  # sorted keys: "a"
    v5 = np.zeros((2, 2, 2, 3, 2))
    v5[..., 0] = 1.0
    with pytest.warns(UserWarning, match="selecting the first"):
        out = fileio.ensure_ndims("x", v5, 4)
    assert out.shape == (2, 2, 2, 3) and np.all(out == 1.0)
    assert fileio.ensure_ndims("x", np.ones((2, 3)), 4).shape == (2, 3, 1, 1)
    with pytest.raises(ValueError, match="No 4-D array"):
        scipy.io.savemat(tmp_path / "none.mat", {"v": np.ones(3)})
        fileio.load_image(tmp_path / "none.mat", 4)
    with pytest.raises(FileNotFoundError, match="no .par or .xml"):
        (tmp_path / "x.rec").write_bytes(b"")
        fileio.load_image(tmp_path / "x.rec", 4)


def test_suffixes():
    assert fileio.maybe_get_suffix("A.NII.GZ") == ".nii.gz"
    assert fileio.chop_allowed_suffix("dir.x/img.nii.gz") == "dir.x/img"
    assert fileio.maybe_get_suffix("img.txt") is None


def test_parse_cli_regparams_and_deprecations(tmp_path):
    base = ["img.mat", "--T2map", "--TE", "0.01", "--nT2", "40", "--T2Range", "0.01", "2.0"]
    o = cli.parse_cli(base + ["--Reg", "chi2", "--RegParams", "1.02"])
    assert o["Chi2Factor"] == 1.02 and o["NoiseLevel"] is None
    o = cli.parse_cli(base + ["--Reg", "mdp", "--RegParams", "0.5"])
    assert o["NoiseLevel"] == 0.5
    with pytest.warns(UserWarning, match="deprecated"):
        o = cli.parse_cli(base + ["--Reg", "chi2", "--Chi2Factor", "1.03"])
    assert o["Chi2Factor"] == 1.03
    with pytest.raises(ValueError, match="both passed"):
        cli.parse_cli(base + ["--Reg", "chi2", "--Chi2Factor", "1.03", "--RegParams", "1.02"])
    with pytest.raises(AssertionError, match="chi2 factor"):
        cli.parse_cli(base + ["--Reg", "chi2"])
    with pytest.raises(ValueError, match="At least one of --T2map or --T2part"):
        cli.parse_cli(["img.mat", "--Reg", "gcv"])
    # settings file: one argument per line, prefixed with @
    s = tmp_path / "settings.txt"
    s.write_text("\n".join(base + ["--Reg", "lcurve", "--SaveRegParam", "-q"]))
    o = cli.parse_cli(["@" + str(s)])
    assert o["Reg"] == "lcurve" and o["SaveRegParam"] and o["quiet"] and o["T2Range"] == [0.01, 2.0]
    kw = cli._kwargs(o, cli._T2MAP_FIELDS)
    assert kw["T2Range"] == (0.01, 2.0) and "SaveNNLSBasis" not in kw and "Chi2Factor" not in kw


def test_file_infos(tmp_path):
    o = cli.parse_cli(["a/x.mat", "b/y.nii.gz", "skip.txt", "--T2part", "-o", "out"])
    infos = cli.get_file_infos(o, [])
    assert [i["choppedinputfile"] for i in infos] == ["x", "y"]
    assert all(i["outputfolder"] == "out" for i in infos)
    o = cli.parse_cli(["a/x.mat", "b/y.nii.gz", "--T2part"])
    assert [i["outputfolder"] for i in cli.get_file_infos(o, [])] == ["a", "b"]
    with pytest.raises(ValueError, match="Number of mask files"):
        cli.get_file_infos(cli.parse_cli(["a/x.mat", "b/y.mat", "--T2part", "-m", "m.mat"]), [])
    with pytest.raises(ValueError, match="Incorrect number of output"):
        cli.get_file_infos(cli.parse_cli(["a/x.mat", "b/y.mat", "c/z.mat", "--T2part", "-o", "p", "q"]), [])
    with pytest.raises(AssertionError, match="SetFlipAngle"):
        cli.get_file_infos(cli.parse_cli(["a/x.mat", "--T2map", "--B1map", "b.mat", "--SetFlipAngle", "170"]), [])


def test_bet_args_and_mask(tmp_path):
    """clean_bet_args (src/main.jl:692-700) and make_bet_mask (:664-681) with a stand-in BET executable that writes
    <out>_mask.nii.gz as FSL's BET does (FSL is not installed here): the first echo goes in as NIfTI, the mask comes
    back and multiplies the image."""
    import sys
    from decaes_amd import cli

    assert cli.clean_bet_args("-m -n -f 0.25 -R") == ["-m", "-n", "-f", "0.25", "-R"]
    assert cli.clean_bet_args("-n,-f,0.5") == ["-m", "-n", "-f", "0.5"]
    fake = tmp_path / "fakebet"
    fake.write_text(f"#!{sys.executable}\n"
                    "import sys\n"
                    f"sys.path.insert(0, {str(Path(cli.__file__).parents[1])!r})\n"
                    "from decaes_amd import fileio\n"
                    "img = fileio.load_image(sys.argv[1], 3)\n"
                    "assert '-m' in sys.argv[3:]\n"
                    "fileio.write_nifti(sys.argv[2] + '_mask.nii.gz', (img > 0.5).astype(float))\n")
    fake.chmod(0o755)
    img = np.random.default_rng(0).uniform(size=(5, 4, 3))
    mask = cli.make_bet_mask(img, str(fake), cli.clean_bet_args("-n"))
    np.testing.assert_array_equal(mask, (img > 0.5).astype(float))
