"""Command line layer: the reference CLI's flags and file conventions (src/main.jl:7-300, :292-420) on this backend.

    python -m decaes_amd.cli image.nii.gz --T2map --T2part --TE 10e-3 --nT2 40 --T2Range 10e-3 2.0 \\
        --SPWin 10e-3 40e-3 --MPWin 40e-3 2.0 --Reg lcurve -o results/
    python -m decaes_amd.cli @settings.txt          # one argument per line, as the reference reads settings files

For each input file: load the 4D image (MAT, NIfTI, PAR/REC or XML/REC), multiply by the mask file if given, run
T2mapSEcorr (with the B1 map if given) and/or T2partSEcorr, and write <name>.t2dist.mat, <name>.t2maps.mat,
<name>.t2parts.mat and <name>.log into the output folder (MAT v7.3, as the reference's MAT.jl writes them), copying
each @settings file there as <name>.<settings file name>. A failure on one file is logged and the next file is
processed. --bet runs the FSL BET executable given by --betpath, as the reference does.
"""
from __future__ import annotations

import argparse
import logging
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
import warnings
from pathlib import Path

import numpy as np

from . import fileio
from .api import T2mapSEcorr, T2partSEcorr
from .options import T2mapOptions, T2partOptions

_T2MAP_FIELDS = ("MatrixSize", "nTE", "TE", "nT2", "T2Range", "T1", "Threshold", "MinRefAngle", "nRefAngles",
                 "nRefAnglesMin", "Reg", "Chi2Factor", "NoiseLevel", "RefConAngle", "SetFlipAngle", "SaveResidualNorm",
                 "SaveDecayCurve", "SaveRegParam", "SaveNNLSBasis", "legacy", "precision")
_T2PART_FIELDS = ("MatrixSize", "nT2", "T2Range", "SPWin", "MPWin", "Sigmoid", "legacy")


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="decaes", fromfile_prefix_chars="@",
                                description="T2mapSEcorr / T2partSEcorr on the gfx950 backend")
    p.add_argument("input", nargs="*", help="input files (.mat, .nii, .nii.gz)")
    p.add_argument("--mask", "-m", nargs="+", default=[], help="mask files, one per input")
    p.add_argument("--output", "-o", nargs="+", default=[], help="output folder(s)")
    p.add_argument("--T2map", action="store_true", help="compute T2 distributions and maps (.t2dist.mat, .t2maps.mat)")
    p.add_argument("--T2part", action="store_true", help="compute T2 parts (.t2parts.mat)")
    p.add_argument("--quiet", "-q", action="store_true")
    p.add_argument("--dry", action="store_true", help="process without saving results")
    p.add_argument("--legacy", action="store_const", const=True, default=None)
    g = p.add_argument_group("T2map/T2part required parameters")
    g.add_argument("--MatrixSize", nargs=3, type=int)
    g.add_argument("--nTE", type=int)
    g.add_argument("--TE", type=float)
    g.add_argument("--nT2", type=int)
    g.add_argument("--T2Range", nargs=2, type=float)
    g.add_argument("--SPWin", nargs=2, type=float)
    g.add_argument("--MPWin", nargs=2, type=float)
    g.add_argument("--Reg", type=str)
    g.add_argument("--RegParams", nargs="+", type=float, default=[])
    g.add_argument("--Chi2Factor", type=float, help="deprecated: use --RegParams")
    g = p.add_argument_group("T2map/T2part optional parameters")
    g.add_argument("--T1", type=float)
    g.add_argument("--Sigmoid", type=float)
    g.add_argument("--Threshold", type=float)
    g = p.add_argument_group("B1 correction and stimulated echo correction")
    g.add_argument("--B1map", nargs="+", default=[])
    g.add_argument("--nRefAngles", type=int)
    g.add_argument("--nRefAnglesMin", type=int)
    g.add_argument("--MinRefAngle", type=float)
    g.add_argument("--SetFlipAngle", type=float)
    g.add_argument("--RefConAngle", type=float)
    g = p.add_argument_group("Additional save options")
    for flag in ("SaveDecayCurve", "SaveNNLSBasis", "SaveRegParam", "SaveResidualNorm"):
        g.add_argument("--" + flag, action="store_true")
    g = p.add_argument_group("BET arguments")
    g.add_argument("--bet", action="store_true")
    g.add_argument("--betargs", type=str, default="-m -n -f 0.25 -R")
    g.add_argument("--betpath", type=str, default="bet")
    g = p.add_argument_group("This backend")
    g.add_argument("--precision", choices=["fp64", "mixed"], default=None)
    return p


def parse_cli(args) -> dict:
    """parse_cli with the reference's deprecations, checks and RegParams mapping (src/main.jl:422-483)."""
    o = vars(build_parser().parse_args(args))
    if o["legacy"] is not None:
        warnings.warn("The flag --legacy is deprecated and will be removed in future releases.")
    if o["Chi2Factor"] is not None:
        if o["RegParams"]:
            raise ValueError("The flag --RegParams and the deprecated flag --Chi2Factor were both passed; use "
                             "--RegParams only.")
        warnings.warn("The flag --Chi2Factor is deprecated and will be removed in future releases; use --RegParams "
                      "instead.")
        o["RegParams"] = [o["Chi2Factor"]]
    o["Chi2Factor"] = None
    o["NoiseLevel"] = None
    if not (o["T2map"] or o["T2part"]):
        raise ValueError("At least one of --T2map or --T2part must be passed")
    if o["Reg"] == "chi2":
        assert len(o["RegParams"]) == 1, 'Must set chi2 factor via --RegParams when --Reg="chi2"'
        o["Chi2Factor"] = o["RegParams"][0]
    elif o["Reg"] == "mdp":
        assert len(o["RegParams"]) == 1, 'Must set noise level via --RegParams when --Reg="mdp"'
        o["NoiseLevel"] = o["RegParams"][0]
    return o


def _kwargs(o: dict, fields) -> dict:
    kw = {}
    for k in fields:
        v = o.get(k)
        if v is None or (isinstance(v, list) and not v) or v is False:
            continue
        kw[k] = tuple(v) if isinstance(v, list) else v
    return kw


def get_file_infos(o: dict, settings_files) -> list:
    """Input, output folder, mask and B1 map per file (src/main.jl:508-564)."""
    inputs = o["input"]
    assert inputs, "At least one input file is required"
    files = [f for f in inputs if fileio.maybe_get_suffix(f)]
    if not files:
        raise ValueError(f"No valid files were found for processing. Note that currently only "
                         f"{', '.join(fileio.ALLOWED_FILE_SUFFIXES)} file types are supported")
    out = o["output"]
    if not out:
        folders = [str(Path(f).parent) for f in files]
    elif len(out) == len(files):
        folders = list(out)
    elif len(out) == 1:
        folders = out * len(files)
    else:
        raise ValueError(f"Incorrect number of output files passed ({len(out)}); must pass either 1 output folder "
                         f"(all results are stored in this folder), or the same number of output folders as input "
                         f"image files ({len(files)})")
    masks = o["mask"]
    if masks and len(masks) != len(files):
        raise ValueError(f"Number of mask files passed ({len(masks)}) does not equal the number of input image files "
                         f"passed ({len(files)})")
    b1 = o["B1map"]
    if b1:
        if len(b1) != len(files):
            raise ValueError(f"Number of B1 map files passed ({len(b1)}) does not equal the number of input image "
                             f"files passed ({len(files)})")
        assert o["SetFlipAngle"] is None, "Cannot set a fixed flip angle using --SetFlipAngle when passing B1 maps using --B1map"
    return [dict(inputfile=f, outputfolder=folders[i], maskfile=masks[i] if masks else None,
                 B1mapfile=b1[i] if b1 else None, choppedinputfile=fileio.chop_allowed_suffix(Path(f).name),
                 settings=list(settings_files)) for i, f in enumerate(files)]


def clean_bet_args(betargs) -> list:
    """clean_bet_args (src/main.jl:692-700): split on commas and spaces; "-m" (make a binary mask) is always there."""
    if not isinstance(betargs, str):
        return list(betargs)
    args = [a for a in re.split(r"[\s,]+", betargs) if a]
    if "-m" not in args:
        args.insert(0, "-m")
    return args


def make_bet_mask(image3: np.ndarray, betpath: str, betargs: list) -> np.ndarray:
    """make_bet_mask (src/main.jl:664-681): run FSL's BET executable on the image (the first echo) written as NIfTI
    in a temporary folder and load the one mask file it writes (BET appends "_mask" and ".nii.gz" to the output
    name)."""
    with tempfile.TemporaryDirectory() as tmp:
        base = Path(tempfile.mktemp(dir=tmp)).name
        src, dst = Path(tmp) / (base + ".nii"), Path(tmp) / (base + ".bet")
        fileio.write_nifti(src, image3)
        subprocess.run([betpath, str(src), str(dst), *betargs], check=True)
        found = sorted(f for f in os.listdir(tmp) if f.startswith(base + ".bet_mask"))
        assert len(found) == 1, found
        return fileio.load_image(Path(tmp) / found[0], 3)


def run_file(info: dict, o: dict, log: logging.Logger) -> None:
    t0 = time.time()
    log.info(f"Loading input file: {info['inputfile']}")
    image = fileio.load_image(info["inputfile"], 4)
    if info["maskfile"] is not None:
        log.info(f"Applying mask from file: {info['maskfile']}")
        try:
            image *= fileio.load_image(info["maskfile"], 3)[..., None]
        except Exception as e:  # the reference warns and continues unmasked
            log.warning(f"Error while loading mask file: {info['maskfile']}: {e}")
    elif o["bet"]:
        args = clean_bet_args(o["betargs"])
        log.info(f"Making and applying BET mask with args: {' '.join(args)}")
        try:
            image *= make_bet_mask(image[..., 0], o["betpath"], args)[..., None]
        except Exception as e:  # the reference warns and continues unmasked
            log.warning(f"Error while making mask using BET: {e}")
    base = Path(info["outputfolder"]) / info["choppedinputfile"]
    if o["T2map"]:
        opts = T2mapOptions.from_image(image, Silent=o["quiet"], **_kwargs(o, _T2MAP_FIELDS))
        alpha = None
        if info["B1mapfile"] is not None:
            log.info(f"Loading B1 map from file: {info['B1mapfile']}")
            try:
                alpha = fileio.load_image(info["B1mapfile"], 3)
            except Exception as e:
                log.warning(f"Error while loading B1 map file: {info['B1mapfile']}: {e}")
        log.info(f"Running T2mapSEcorr on file: {info['inputfile']}")
        maps, dist = T2mapSEcorr(image, opts, alpha=alpha)
        if not o["dry"]:
            log.info(f"Saving T2 distribution to file: {base}.t2dist.mat")
            fileio.save_mat(f"{base}.t2dist.mat", {"dist": dist})
            log.info(f"Saving T2 parameter maps to file: {base}.t2maps.mat")
            fileio.save_mat(f"{base}.t2maps.mat", maps)
    else:
        dist = image
    if o["T2part"]:
        kw = _kwargs(o, _T2PART_FIELDS)
        if not o["T2map"]:
            kw.pop("nT2", None)  # inferred from the distribution (src/main.jl t2part_options)
        popts = T2partOptions.from_dist(dist, Silent=o["quiet"], **kw)
        log.info("Running T2partSEcorr")
        parts = T2partSEcorr(dist, popts)
        if not o["dry"]:
            log.info(f"Saving T2 parts maps to file: {base}.t2parts.mat")
            fileio.save_mat(f"{base}.t2parts.mat", parts)
    log.info(f"Finished ({time.time() - t0:.2f} seconds)")


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    o = parse_cli(argv)
    settings_files = [a[1:] for a in argv if a.startswith("@")]
    status = 0
    for info in get_file_infos(o, settings_files):
        folder = Path(info["outputfolder"])
        if not o["dry"]:
            folder.mkdir(parents=True, exist_ok=True)
            for s in info["settings"]:
                shutil.copyfile(s, folder / f"{info['choppedinputfile']}.{Path(s).name}")
        log = logging.getLogger(f"decaes.{info['choppedinputfile']}.{id(info)}")
        log.setLevel(logging.INFO)
        log.propagate = False
        handlers = []
        if not o["quiet"]:
            handlers.append(logging.StreamHandler(sys.stdout))
        if not o["dry"]:
            handlers.append(logging.FileHandler(folder / f"{info['choppedinputfile']}.log", mode="w"))
        for h in handlers:
            h.setFormatter(logging.Formatter("[%(asctime)s] %(levelname)s %(message)s"))
            log.addHandler(h)
        try:
            run_file(info, o, log)
        except Exception as e:  # one failing file does not stop the others (src/main.jl:329-336)
            log.warning(f"Error during processing of file: {info['inputfile']}: {type(e).__name__}: {e}")
            status = 1
        finally:
            for h in handlers:
                h.close()
                log.removeHandler(h)
    return status


if __name__ == "__main__":
    sys.exit(main())
