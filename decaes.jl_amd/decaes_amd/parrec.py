"""Philips PAR/REC and XML/REC images: the third input format of load_image (src/main.jl:601-603), which the
reference reads with ParXRec.load from the ParXRec.jl package (version 0.2.0 in the reference's manifest,
test/interactive/pluto/interactive_notebook.jl:1027-1031) and unwraps to its plain array (`parent(rec.data)`).

ParXRec.jl is not part of the reference tree and the reference's tests hold no PAR, XML or REC file
(test/cli.jl:290), so this reader follows the published file formats, not that package's code:

* PAR (versions 4, 4.1, 4.2): a text header whose "IMAGE INFORMATION DEFINITION" block names the columns of the
  image lines (with their counts, e.g. "recon resolution (x y)  (2*integer)"), followed by one line per image.
* XML (Philips PRIDE XML): one `Image_Info` element per image, its `Key` (Slice, Echo, Dynamic, Phase, Type,
  Sequence, Index) and its attributes (Pixel Size, Resolution X/Y, Rescale Intercept/Slope, Scale Slope) given as
  `<Attribute Name="...">value</Attribute>` elements.
* REC: the images as little-endian unsigned integers of the stated pixel size (8 or 16 bits), x fastest within an
  image, image k at byte offset k * X * Y * bytes, k the image's "index in REC file" (the line order where a PAR
  version 4 file has no index column).

Each image's pixel values PV are scaled to Philips' floating-point value FP = (PV * RS + RI) / (RS * SS) (RS the
rescale slope, RI the rescale intercept, SS the scale slope) by default; `scaling="dv"` gives the displayed value
PV * RS + RI and `scaling="pv"` the stored integers. The array's axes are (x, y, slice, echo) followed by dynamic,
cardiac phase, image type and scanning sequence where these have more than one value, each axis in increasing order
of its label, which puts the echoes of a multi-echo scan on the fourth axis as load_image expects.

Parity with ParXRec.jl is unpinned: its default scaling and axis order are not in any file the reference holds.
"""
from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

_PAR_FIELDS = {
    "slice": "slice number",
    "echo": "echo number",
    "dynamic": "dynamic scan number",
    "phase": "cardiac phase number",
    "type": "image_type_mr",
    "sequence": "scanning sequence",
    "index": "index in rec file",
    "bits": "image pixel size",
    "recon": "recon resolution",
    "ri": "rescale intercept",
    "rs": "rescale slope",
    "ss": "scale slope",
}
_XML_FIELDS = {
    "slice": "Slice", "echo": "Echo", "dynamic": "Dynamic", "phase": "Phase", "type": "Type", "sequence": "Sequence",
    "index": "Index", "bits": "Pixel Size", "rx": "Resolution X", "ry": "Resolution Y", "ri": "Rescale Intercept",
    "rs": "Rescale Slope", "ss": "Scale Slope",
}
_AXES = ("slice", "echo", "dynamic", "phase", "type", "sequence")
_DEF_LINE = re.compile(r"^#\s+(.*\S)\s+\((?:(\d+)\*)?(integer|float|string)\)\s*$")


def _companion(path: Path, suffixes) -> Path:
    for s in suffixes:
        for cand in (path.with_suffix(s.lower()), path.with_suffix(s.upper())):
            if cand.exists():
                return cand
    raise FileNotFoundError(f"{path}: no {' or '.join(suffixes)} file beside it")


def _number(tok: str):
    try:
        return int(tok)
    except ValueError:
        return float(tok)


def read_par(path) -> list:
    """The image lines of a PAR file (version 4 or later) as dicts with the keys of _PAR_FIELDS."""
    path = Path(path)
    lines = path.read_text(encoding="latin-1").splitlines()
    columns = []  # (name, count, type) in line order
    section = None
    images = []
    for raw in lines:
        line = raw.strip()
        if line.startswith("#"):
            up = line.upper()
            if "IMAGE INFORMATION DEFINITION" in up:
                section = "def"
            elif "IMAGE INFORMATION" in up:
                section = "img"
            elif "END OF DATA DESCRIPTION" in up:
                section = None
            elif section == "def":
                m = _DEF_LINE.match(line)
                if m:
                    columns.append((m.group(1).lower(), int(m.group(2) or 1), m.group(3)))
            continue
        if section != "img" or not line:
            continue
        if not columns:
            raise ValueError(f"{path}: no IMAGE INFORMATION DEFINITION block (PAR versions before 4 are not read)")
        toks = line.split()
        need = sum(c for _, c, _ in columns)
        if len(toks) < need:
            raise ValueError(f"{path}: an image line has {len(toks)} values, the definition block names {need}")
        row, pos = {}, 0
        for name, count, kind in columns:
            vals = [t if kind == "string" else _number(t) for t in toks[pos:pos + count]]
            row[name] = vals if count > 1 else vals[0]
            pos += count
        images.append(row)
    if not images:
        raise ValueError(f"{path}: no image lines")
    out = []
    for k, row in enumerate(images):
        img = {}
        for key, prefix in _PAR_FIELDS.items():
            hit = [v for n, v in row.items() if n.startswith(prefix)]
            if hit:
                img[key] = hit[0]
            elif key == "index":
                img[key] = k
            elif key in ("dynamic", "phase", "type", "sequence"):
                img[key] = 1 if key in ("dynamic", "phase") else 0
            else:
                raise ValueError(f"{path}: the image lines have no '{prefix}' column")
        img["rx"], img["ry"] = (int(v) for v in img.pop("recon"))
        out.append(img)
    return out


def read_xml(path) -> list:
    """The Image_Info elements of a Philips XML header as dicts with the keys of _XML_FIELDS."""
    path = Path(path)
    root = ET.parse(path).getroot()
    out = []
    for info in root.iter("Image_Info"):
        attrs = {}
        for a in info.iter("Attribute"):
            attrs[a.get("Name")] = (a.text or "").strip()
        img = {}
        for key, name in _XML_FIELDS.items():
            if name not in attrs:
                if key in ("dynamic", "phase", "type", "sequence"):
                    img[key] = 1 if key in ("dynamic", "phase") else ""
                    continue
                raise ValueError(f"{path}: an Image_Info has no '{name}' attribute")
            v = attrs[name]
            img[key] = v if key in ("type", "sequence") else _number(v)
        out.append(img)
    if not out:
        raise ValueError(f"{path}: no Image_Info elements")
    return out


def load(filename, scaling: str = "fp") -> np.ndarray:
    """The image of a PAR/REC or XML/REC pair, given any one of its files, as a float64 array (x, y, slice, echo,
    ...) in column-major order."""
    if scaling not in ("fp", "dv", "pv"):
        raise ValueError(f"scaling must be 'fp', 'dv' or 'pv', not {scaling!r}")
    path = Path(filename)
    ext = path.suffix.lower()
    if ext == ".rec":
        header = _companion(path, (".par", ".xml"))
        rec = path
    elif ext in (".par", ".xml"):
        header = path
        rec = _companion(path, (".rec",))
    else:
        raise ValueError(f"{filename}: not a .par, .xml or .rec file")
    images = read_xml(header) if header.suffix.lower() == ".xml" else read_par(header)

    rx, ry, bits = images[0]["rx"], images[0]["ry"], images[0]["bits"]
    if any((im["rx"], im["ry"], im["bits"]) != (rx, ry, bits) for im in images):
        raise ValueError(f"{header}: images of different sizes or pixel sizes")
    if bits not in (8, 16):
        raise ValueError(f"{header}: pixel size {bits} bits (8 or 16 are read)")
    dtype = np.dtype("<u1" if bits == 8 else "<u2")
    raw = np.fromfile(rec, dtype=dtype)
    per = rx * ry
    nimg = raw.size // per
    labels = {ax: sorted({im[ax] for im in images}) for ax in _AXES}
    shape = tuple(len(labels[ax]) for ax in _AXES)
    pos = {ax: {v: i for i, v in enumerate(labels[ax])} for ax in _AXES}
    out = np.full((rx, ry) + shape, np.nan, order="F")
    seen = np.zeros(shape, dtype=bool)
    for im in images:
        k = int(im["index"])
        if not 0 <= k < nimg:
            raise ValueError(f"{rec}: image index {k} beyond the {nimg} images the file holds")
        where = tuple(pos[ax][im[ax]] for ax in _AXES)
        if seen[where]:
            raise ValueError(f"{header}: two images share slice/echo/dynamic/phase/type/sequence {where}")
        seen[where] = True
        pv = raw[k * per:(k + 1) * per].astype(np.float64).reshape((rx, ry), order="F")
        rs, ri, ss = float(im["rs"]), float(im["ri"]), float(im["ss"])
        if scaling == "pv":
            val = pv
        elif scaling == "dv":
            val = pv * rs + ri
        else:
            val = (pv * rs + ri) / (rs * ss)
        out[(slice(None), slice(None)) + where] = val
    if not seen.all():
        raise ValueError(f"{header}: {int((~seen).sum())} slice/echo/dynamic/phase/type/sequence combinations "
                         "have no image")
    # (x, y, slice, echo) always; dynamic, phase, type and sequence only where they have more than one value
    kept = tuple(out.shape[:4]) + tuple(d for d in out.shape[4:] if d > 1)
    return out.reshape(kept, order="F")
