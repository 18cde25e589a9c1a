"""MAT v7.3 output and input (decaes_amd.mat73), the container the reference's MAT.jl writes results in
(src/main.jl:382,391,411) and reads images from (src/main.jl:577-589).

Checked against the HDF5 project's own h5dump (an independent reader of the same file) where it is installed: the
dataset holds float64 with the dimensions reversed and the column-major bytes unchanged, and carries MATLAB's
MATLAB_class attribute; the 512-byte user block holds the MAT v7.3 header."""
import shutil
import subprocess

import numpy as np
import pytest

from decaes_amd import fileio, mat73

pytestmark = pytest.mark.skipif(not mat73.available(), reason="libhdf5 not installed")

H5DUMP = shutil.which("h5dump") or (shutil.which("h5dump", path="/opt/conda/bin"))


def _maps(rng):
    gdn = rng.uniform(size=(5, 4, 3))
    gdn[1, 2, 0] = np.nan
    return {"echotimes": 0.01 * np.arange(1, 33), "t2times": np.logspace(-2, np.log10(2.0), 40),
            "gdn": gdn, "decaycurve": rng.normal(size=(5, 4, 3, 32)),
            "decaybasis": rng.normal(size=(2, 3, 2, 7, 5)), "decaybasisset": rng.normal(size=(32, 40, 8))}


def test_roundtrip_is_bitwise(tmp_path):
    arrays = _maps(np.random.default_rng(0))
    mat73.write(tmp_path / "x.mat", arrays)
    back = mat73.read(tmp_path / "x.mat")
    assert set(back) == set(arrays)
    for k, v in arrays.items():
        assert back[k].shape == v.shape, k
        assert np.array_equal(back[k], v, equal_nan=True), k


def test_header_and_superblock(tmp_path):
    mat73.write(tmp_path / "x.mat", {"a": np.ones((2, 2))})
    raw = (tmp_path / "x.mat").read_bytes()
    assert raw[:19] == b"MATLAB 7.3 MAT-file"
    assert raw[125:128] == b"\x02IM"
    assert raw[512:520] == b"\x89HDF\r\n\x1a\n"  # the HDF5 superblock after the 512-byte user block
    assert mat73.is_mat73(tmp_path / "x.mat")


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_layout_seen_by_h5dump(tmp_path):
    a = np.arange(24.0).reshape((4, 3, 2), order="F")
    mat73.write(tmp_path / "x.mat", {"gdn": a, "v": np.arange(3.0)})
    out = subprocess.run([H5DUMP, "-d", "/gdn", str(tmp_path / "x.mat")], capture_output=True, text=True,
                         check=True).stdout
    assert "H5T_IEEE_F64LE" in out
    assert "SIMPLE { ( 2, 3, 4 ) / ( 2, 3, 4 ) }" in out
    assert "(0,0,0): 0, 1, 2, 3," in out  # the first column of the column-major array is the fastest HDF5 axis
    assert 'ATTRIBUTE "MATLAB_class"' in out and "STRSIZE 6;" in out and "H5T_STR_NULLTERM" in out
    assert "H5T_CSET_ASCII" in out and "SCALAR" in out and '"double"' in out
    out = subprocess.run([H5DUMP, "-H", "-d", "/v", str(tmp_path / "x.mat")], capture_output=True, text=True,
                         check=True).stdout
    assert "SIMPLE { ( 3 ) / ( 3 ) }" in out


def test_load_image_and_save_mat(tmp_path):
    rng = np.random.default_rng(1)
    img4 = rng.uniform(size=(4, 3, 2, 5))
    mat73.write(tmp_path / "in.mat", {"zimg": img4, "b": img4 * 2, "m3": np.ones((4, 3, 2))})
    with pytest.warns(UserWarning):
        np.testing.assert_array_equal(fileio.load_image(tmp_path / "in.mat", 4), img4 * 2)  # sorted keys: "b"
    np.testing.assert_array_equal(fileio.load_image(tmp_path / "in.mat", 3), np.ones((4, 3, 2)))
    fileio.save_mat(tmp_path / "out.mat", {"dist": img4})
    assert mat73.is_mat73(tmp_path / "out.mat")
    np.testing.assert_array_equal(fileio.read_mat(tmp_path / "out.mat")["dist"], img4)
