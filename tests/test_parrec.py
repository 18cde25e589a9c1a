"""PAR/REC and XML/REC input (decaes_amd.parrec; load_image, src/main.jl:601-603, which reads these files with the
third-party ParXRec.jl 0.2.0). The reference's tests hold no such file (test/cli.jl:290), so parity with ParXRec.jl
is unpinned; these tests hold the reader to the published formats: files laid out as the scanner's export writes them
(tests/parrec_files.py), image lines and REC images in shuffled orders, per-image scaling, and the expected array
built from the known pixel values by slice, echo and dynamic number."""
import numpy as np
import pytest

import parrec_files as PF
from decaes_amd import fileio, parrec


def _pixels(shape, seed=0):
    return np.random.default_rng(seed).integers(0, 4096, size=shape).astype(np.uint16)


def _scales(shape, seed=1):
    rng = np.random.default_rng(seed)
    nz, ne, nd = shape[2:]
    return (rng.uniform(0.5, 4.0, (nz, ne, nd)), rng.uniform(-5.0, 5.0, (nz, ne, nd)),
            rng.uniform(1e-3, 1e-2, (nz, ne, nd)))


def _expected(pv, rs, ri, ss):
    return (pv * rs[None, None] + ri[None, None]) / (rs[None, None] * ss[None, None])


@pytest.mark.parametrize("version", ["4.2", "4"])
def test_par_rec(tmp_path, version):
    shape = (5, 4, 3, 6, 2)  # x, y, slice, echo, dynamic
    pv = _pixels(shape)
    rs, ri, ss = _scales(shape)
    rec = PF.write_par(tmp_path / "scan.PAR", pv, rs, ri, ss, version=version)
    PF.write_rec(tmp_path / "scan.REC", rec)
    assert rec.shape == (3 * 6 * 2, 20)
    a = parrec.load(tmp_path / "scan.PAR")
    assert a.shape == shape and a.dtype == np.float64
    np.testing.assert_allclose(a, _expected(pv, rs, ri, ss), rtol=1e-15, atol=0)
    # the .rec file of the pair names the same image
    np.testing.assert_array_equal(parrec.load(tmp_path / "scan.REC"), a)
    np.testing.assert_array_equal(parrec.load(tmp_path / "scan.PAR", scaling="pv"), pv)
    np.testing.assert_allclose(parrec.load(tmp_path / "scan.PAR", scaling="dv"), pv * rs[None, None] + ri[None, None],
                               rtol=1e-15)
    # load_image keeps (x, y, slice, echo) and takes the first dynamic, with the reference's warning
    with pytest.warns(UserWarning, match="selecting the first"):
        img = fileio.load_image(tmp_path / "scan.PAR", 4)
    np.testing.assert_array_equal(img, a[..., 0])
    assert img.flags.f_contiguous


def test_xml_rec_equals_par_rec(tmp_path):
    shape = (6, 3, 2, 5, 1)
    pv = _pixels(shape, seed=4)
    rs, ri, ss = _scales(shape, seed=5)
    PF.write_rec(tmp_path / "a.rec", PF.write_par(tmp_path / "a.par", pv, rs, ri, ss, seed=7))
    PF.write_rec(tmp_path / "b.rec", PF.write_xml(tmp_path / "b.xml", pv, rs, ri, ss, seed=8))
    a = parrec.load(tmp_path / "a.par")
    b = parrec.load(tmp_path / "b.xml")
    assert a.shape == b.shape == shape[:4]  # a single dynamic adds no axis
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(b, _expected(pv, rs, ri, ss)[..., 0], rtol=1e-15)
    np.testing.assert_array_equal(fileio.load_image(tmp_path / "b.xml", 4), b)
    np.testing.assert_array_equal(fileio.load_image(tmp_path / "b.rec", 4), b)


def test_errors(tmp_path):
    shape = (4, 4, 2, 3, 1)
    pv = _pixels(shape)
    rec = PF.write_par(tmp_path / "s.par", pv, 1.0, 0.0, 1.0)
    with pytest.raises(FileNotFoundError, match="no .rec"):
        parrec.load(tmp_path / "s.par")
    PF.write_rec(tmp_path / "s.rec", rec[:-1])  # one image short
    with pytest.raises(ValueError, match="beyond the 5 images"):
        parrec.load(tmp_path / "s.par")
    # a header that lists one image twice leaves another combination without an image
    text = (tmp_path / "s.par").read_text(encoding="latin-1").splitlines()
    img_lines = [i for i, l in enumerate(text) if l.strip() and not l.lstrip().startswith(("#", "."))]
    toks = text[img_lines[1]].split()
    toks[:3] = text[img_lines[0]].split()[:3]
    text[img_lines[1]] = " " + " ".join(toks)
    (tmp_path / "t.par").write_text("\n".join(text), encoding="latin-1")
    PF.write_rec(tmp_path / "t.rec", rec)
    with pytest.raises(ValueError, match="share slice/echo"):
        parrec.load(tmp_path / "t.par")
    (tmp_path / "u.par").write_text("# === IMAGE INFORMATION ===\n 1 1 1\n", encoding="latin-1")
    PF.write_rec(tmp_path / "u.rec", rec)
    with pytest.raises(ValueError, match="DEFINITION"):
        parrec.load(tmp_path / "u.par")
    with pytest.raises(ValueError, match="scaling"):
        parrec.load(tmp_path / "s.par", scaling="raw")
