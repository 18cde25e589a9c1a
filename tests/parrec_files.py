"""Writers of Philips PAR/REC (version 4 and 4.2) and XML/REC test files, laid out as the scanner's export writes
them: the PAR header's general information, pixel-value and image-definition blocks, one image line per image; the
XML header's Series_Info and one Image_Info per image; the REC file as little-endian unsigned 16-bit images, x fastest.
Used by the reader's tests only."""
from __future__ import annotations

import itertools
from pathlib import Path

import numpy as np

PAR_HEAD = """# === DATA DESCRIPTION FILE ======================================================
#
# CAUTION - Investigational device.
# Limited by Federal Law to investigational use.
#
# Dataset name: D:\\\\Export\\\\{name}
#
# CLINICAL TRYOUT             Research image export tool     V{version}
#
# === GENERAL INFORMATION ========================================================
#
.    Patient name                       :   phantom
.    Examination name                   :   relaxometry
.    Protocol name                      :   WIP MSE 32 echoes
.    Examination date/time              :   2024.01.01 / 12:00:00
.    Series Type                        :   Image   MRSERIES
.    Acquisition nr                     :   3
.    Reconstruction nr                  :   1
.    Scan Duration [sec]                :   612
.    Max. number of cardiac phases      :   1
.    Max. number of echoes              :   {ne}
.    Max. number of slices/locations    :   {nz}
.    Max. number of dynamics            :   {nd}
.    Max. number of mixes               :   1
.    Patient position                   :   Head First Supine
.    Preparation direction              :   Right-Left
.    Technique                          :   SE
.    Scan resolution  (x, y)            :   {nx}  {ny}
.    Scan mode                          :   2D
.    Repetition time [ms]               :   1200.000
.    FOV (ap,fh,rl) [mm]                :   {ny}.000  {nz}.000  {nx}.000
.    Water Fat shift [pixels]           :   1.000
.    Angulation midslice(ap,fh,rl)[degr]:   0.000  0.000  0.000
.    Off Centre midslice(ap,fh,rl) [mm] :   0.000  0.000  0.000
.    Flow compensation <0=no 1=yes> ?   :   0
.    Presaturation     <0=no 1=yes> ?   :   0
.    Phase encoding velocity [cm/sec]   :   0.000000  0.000000  0.000000
.    MTC               <0=no 1=yes> ?   :   0
.    SPIR              <0=no 1=yes> ?   :   0
.    EPI factor        <0,1=no EPI>     :   1
.    Dynamic scan      <0=no 1=yes> ?   :   {dyn}
.    Diffusion         <0=no 1=yes> ?   :   0
.    Diffusion echo time [ms]           :   0.0000
.    Max. number of diffusion values    :   1
.    Max. number of gradient orients    :   1
.    Number of label types   <0=no ASL> :   0
#
# === PIXEL VALUES =============================================================
#  PV = pixel value in REC file, FP = floating point value, DV = displayed value on console
#  RS = rescale slope,           RI = rescale intercept,    SS = scale slope
#  DV = PV * RS + RI             FP = DV / (RS * SS)
#
# === IMAGE INFORMATION DEFINITION =============================================
#  The rest of this file contains ONE line per image, this line contains the following information:
#
"""

# (name, count, kind) of the version 4.2 image line
COLUMNS_42 = [
    ("slice number", 1, "integer"), ("echo number", 1, "integer"), ("dynamic scan number", 1, "integer"),
    ("cardiac phase number", 1, "integer"), ("image_type_mr", 1, "integer"), ("scanning sequence", 1, "integer"),
    ("index in REC file (in images)", 1, "integer"), ("image pixel size (in bits)", 1, "integer"),
    ("scan percentage", 1, "integer"), ("recon resolution (x y)", 2, "integer"), ("rescale intercept", 1, "float"),
    ("rescale slope", 1, "float"), ("scale slope", 1, "float"), ("window center", 1, "integer"),
    ("window width", 1, "integer"), ("image angulation (ap,fh,rl in degrees )", 3, "float"),
    ("image offcentre (ap,fh,rl in mm )", 3, "float"), ("slice thickness (in mm )", 1, "float"),
    ("slice gap (in mm )", 1, "float"), ("image_display_orientation", 1, "integer"),
    ("slice orientation ( TRA/SAG/COR )", 1, "integer"), ("fmri_status_indication", 1, "integer"),
    ("image_type_ed_es  (end diast/end syst)", 1, "integer"), ("pixel spacing (x,y) (in mm)", 2, "float"),
    ("echo_time", 1, "float"), ("dyn_scan_begin_time", 1, "float"), ("trigger_time", 1, "float"),
    ("diffusion_b_factor", 1, "float"), ("number of averages", 1, "integer"),
    ("image_flip_angle (in degrees)", 1, "float"), ("cardiac frequency   (bpm)", 1, "integer"),
    ("minimum RR-interval (in ms)", 1, "integer"), ("maximum RR-interval (in ms)", 1, "integer"),
    ("TURBO factor  <0=no turbo>", 1, "integer"), ("Inversion delay (in ms)", 1, "float"),
    ("diffusion b value number    (imagekey!)", 1, "integer"),
    ("gradient orientation number (imagekey!)", 1, "integer"), ("contrast type", 1, "string"),
    ("diffusion anisotropy type", 1, "string"), ("diffusion (ap, fh, rl)", 3, "float"),
    ("label type (ASL)            (imagekey!)", 1, "integer"),
]
# version 4: no index column (images are stored in line order) and none of the later additions
COLUMNS_4 = [c for c in COLUMNS_42[:35] if not c[0].startswith("index in REC")]


def _images(shape, rng, order):
    """(slice, echo, dynamic) of each image and its position in the REC file."""
    nz, ne, nd = shape[2], shape[3], shape[4]
    keys = list(itertools.product(range(nz), range(ne), range(nd)))
    pos = np.arange(len(keys)) if order == "lines" else rng.permutation(len(keys))
    return keys, pos


def write_par(path, pv: np.ndarray, rs, ri, ss, version="4.2", seed=0, te=0.01):
    """pv: uint16 (x, y, slice, echo, dynamic); rs, ri, ss: arrays of shape (slice, echo, dynamic) or scalars.
    Image lines are written echo-major in a shuffled line order, and (version 4.2) the images are stored in the REC
    file in another shuffled order given by the index column."""
    path = Path(path)
    rng = np.random.default_rng(seed)
    nx, ny, nz, ne, nd = pv.shape
    rs, ri, ss = (np.broadcast_to(np.asarray(v, dtype=float), (nz, ne, nd)) for v in (rs, ri, ss))
    cols = COLUMNS_42 if version == "4.2" else COLUMNS_4
    keys, pos = _images(pv.shape, rng, "shuffled" if version == "4.2" else "lines")
    lines = rng.permutation(len(keys)) if version == "4.2" else np.arange(len(keys))
    text = PAR_HEAD.format(name=path.stem, version=version, ne=ne, nz=nz, nd=nd, nx=nx, ny=ny, dyn=int(nd > 1))
    for name, count, kind in cols:
        kind_s = f"{count}*{kind}" if count > 1 else kind
        text += f"#  {name:<40} ({kind_s})\n"
    text += "#\n# === IMAGE INFORMATION ==========================================================\n"
    text += "#  sl ec  dyn ph ty    idx pix scan% rec size                (re)scale              window\n\n"
    rec = np.zeros((len(keys), nx * ny), dtype="<u2")
    for li in lines:
        z, e, d = keys[li]
        rec[pos[li]] = pv[:, :, z, e, d].reshape(-1, order="F")
        vals = {"slice number": [z + 1], "echo number": [e + 1], "dynamic scan number": [d + 1],
                "cardiac phase number": [1], "image_type_mr": [0], "scanning sequence": [2],
                "index in REC file (in images)": [int(pos[li])], "image pixel size (in bits)": [16],
                "scan percentage": [100], "recon resolution (x y)": [nx, ny],
                "rescale intercept": [f"{ri[z, e, d]:.17g}"], "rescale slope": [f"{rs[z, e, d]:.17g}"],
                "scale slope": [f"{ss[z, e, d]:.17g}"], "echo_time": [f"{1000 * te * (e + 1):.2f}"],
                "image_flip_angle (in degrees)": ["90.00"], "contrast type": ["T2"],
                "diffusion anisotropy type": ["-"]}
        toks = []
        for name, count, kind in cols:
            v = vals.get(name)
            if v is None:
                v = ["0.000" if kind == "float" else "0"] * count
            toks += [str(t) for t in v]
        text += " " + " ".join(toks) + "\n"
    text += "\n# === END OF DATA DESCRIPTION FILE ===============================================\n"
    path.write_text(text, encoding="latin-1")
    return rec


def write_xml(path, pv: np.ndarray, rs, ri, ss, seed=0):
    """The same image as a Philips XML header (PRIDE V5) with shuffled Image_Info order and REC order."""
    path = Path(path)
    rng = np.random.default_rng(seed)
    nx, ny, nz, ne, nd = pv.shape
    rs, ri, ss = (np.broadcast_to(np.asarray(v, dtype=float), (nz, ne, nd)) for v in (rs, ri, ss))
    keys, pos = _images(pv.shape, rng, "shuffled")
    rec = np.zeros((len(keys), nx * ny), dtype="<u2")
    out = ['<?xml version="1.0" encoding="UTF-8"?>', "<PRIDE_V5>", "<Series_Info>",
           '<Attribute Name="Patient Name" Type="String">phantom</Attribute>',
           f'<Attribute Name="Max No Echoes" Type="Int32">{ne}</Attribute>', "</Series_Info>", "<Image_Array>"]
    for li in rng.permutation(len(keys)):
        z, e, d = keys[li]
        rec[pos[li]] = pv[:, :, z, e, d].reshape(-1, order="F")
        out += ["<Image_Info>", "<Key>",
                f'<Attribute Name="Slice" Type="Int32">{z + 1}</Attribute>',
                f'<Attribute Name="Echo" Type="Int32">{e + 1}</Attribute>',
                f'<Attribute Name="Dynamic" Type="Int32">{d + 1}</Attribute>',
                '<Attribute Name="Phase" Type="Int32">1</Attribute>',
                '<Attribute Name="BValue" Type="Int32">1</Attribute>',
                '<Attribute Name="Grad Orient" Type="Int32">1</Attribute>',
                '<Attribute Name="Label Type" Type="Enumeration" EnumType="Label_Type">-</Attribute>',
                '<Attribute Name="Type" Type="Enumeration" EnumType="Type">M</Attribute>',
                '<Attribute Name="Sequence" Type="Enumeration" EnumType="Sequence">SE</Attribute>',
                f'<Attribute Name="Index" Type="Int32">{int(pos[li])}</Attribute>', "</Key>",
                '<Attribute Name="Pixel Size" Type="UInt16">16</Attribute>',
                '<Attribute Name="Scan Percentage" Type="Double">100</Attribute>',
                f'<Attribute Name="Resolution X" Type="UInt16">{nx}</Attribute>',
                f'<Attribute Name="Resolution Y" Type="UInt16">{ny}</Attribute>',
                f'<Attribute Name="Rescale Intercept" Type="Double">{ri[z, e, d]:.17g}</Attribute>',
                f'<Attribute Name="Rescale Slope" Type="Double">{rs[z, e, d]:.17g}</Attribute>',
                f'<Attribute Name="Scale Slope" Type="Double">{ss[z, e, d]:.17g}</Attribute>',
                "</Image_Info>"]
    out += ["</Image_Array>", "</PRIDE_V5>"]
    path.write_text("\n".join(out) + "\n")
    return rec


def write_rec(path, rec: np.ndarray):
    Path(path).write_bytes(np.ascontiguousarray(rec, dtype="<u2").tobytes())
