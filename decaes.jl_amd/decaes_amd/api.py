"""T2mapSEcorr / T2partSEcorr on the gfx950 backend, with the reference's signatures and outputs.

`T2mapSEcorr(image, opts)` returns `(maps, dist)` where `maps` holds exactly the keys the reference's
`convert(Dict{String,Any}, maps)` keeps (src/T2mapSEcorr.jl:21-54) and `dist` is the (X,Y,Z,nT2)
distribution array; `T2partSEcorr(dist, opts)` returns {"sfr","sgm","mfr","mgm"} (src/T2partSEcorr.jl:40-93).
Arrays are Fortran-ordered, as Julia's column-major arrays. Float64 inputs run the FP64 path; Float32 inputs (the
reference's T{Float32} methods) return Float32 outputs: the FP64 path's results rounded once. Every voxel runs on
the device.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from typing import Optional

import numpy as np

from . import _lib
from .options import T2mapOptions, T2partOptions

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_fp if a.dtype == np.float32 else _dp)


def _nan(shape, dtype=np.float64):
    return np.full(shape, np.nan, order="F", dtype=dtype)


def _float_type(a) -> type:
    return np.float32 if np.asarray(a).dtype == np.float32 else np.float64


def t2map_opts_c(o: T2mapOptions) -> _lib.T2mapOptsC:
    c = _lib.T2mapOptsC()
    for k in range(3):
        c.MatrixSize[k] = o.MatrixSize[k]
    c.nTE, c.nT2, c.TE = o.nTE, o.nT2, o.TE
    c.T2Range[0], c.T2Range[1] = o.T2Range
    c.T1, c.Threshold, c.MinRefAngle = o.T1, o.Threshold, o.MinRefAngle
    c.nRefAngles, c.nRefAnglesMin = o.nRefAngles, o.nRefAnglesMin
    c.Reg = _lib.REG[o.Reg]
    c.legacy = int(o.legacy)
    c.Chi2Factor = math.nan if o.Chi2Factor is None else o.Chi2Factor
    c.NoiseLevel = math.nan if o.NoiseLevel is None else o.NoiseLevel
    c.RefConAngle = o.RefConAngle
    c.SetFlipAngle = math.nan if o.SetFlipAngle is None else o.SetFlipAngle
    c.SaveResidualNorm, c.SaveDecayCurve = int(o.SaveResidualNorm), int(o.SaveDecayCurve)
    c.SaveRegParam, c.SaveNNLSBasis = int(o.SaveRegParam), int(o.SaveNNLSBasis)
    c.precision = _lib.PREC[o.precision]
    c.schedule = _lib.SCHED[o.schedule]
    return c


def t2part_opts_c(o: T2partOptions) -> _lib.T2partOptsC:
    c = _lib.T2partOptsC()
    for k in range(3):
        c.MatrixSize[k] = o.MatrixSize[k]
    c.nT2 = o.nT2
    c.legacy = int(o.legacy)
    c.T2Range[0], c.T2Range[1] = o.T2Range
    c.SPWin[0], c.SPWin[1] = o.SPWin
    c.MPWin[0], c.MPWin[1] = o.MPWin
    c.Sigmoid = math.nan if o.Sigmoid is None else o.Sigmoid
    return c


def decaybasisset(opts: T2mapOptions, device: Optional[int] = None) -> np.ndarray:
    """maps["decaybasisset"]: nTE x nT2 x nRefAngles grid bases (nTE x nT2 with SetFlipAngle)."""
    ctx = _lib.context(device)
    K = opts.nRefAngles if opts.SetFlipAngle is None else 1
    shape = (opts.nTE, opts.nT2, K) if K > 1 or opts.SetFlipAngle is None else (opts.nTE, opts.nT2)
    out = np.empty(shape, order="F")
    oc = t2map_opts_c(opts)
    _lib.check(_lib.load().dmi_epg_basis_set(ctx.handle, C.byref(oc), _ptr(out)), ctx.handle)
    return out


def T2mapSEcorr(image: np.ndarray, opts: Optional[T2mapOptions] = None, *, alpha: Optional[np.ndarray] = None,
                device: Optional[int] = None, **kwargs):
    """T2mapSEcorr(image, opts) (src/T2mapSEcorr.jl:148-196) -> (maps, dist).

    `alpha`, if given, is a B1 (flip angle) map as passed through load_B1map! (src/T2mapSEcorr.jl:56-61):
    the flip angle search is skipped and the map is echoed in maps["alpha"].
    """
    ft = _float_type(image)
    image = np.asfortranarray(image, dtype=ft)
    if opts is None:
        opts = T2mapOptions.from_image(image, **kwargs)
    elif kwargs:
        raise TypeError("pass either opts or keyword options, not both")
    assert image.shape == (*opts.MatrixSize, opts.nTE)
    ctx = _lib.context(device)
    shape3 = opts.MatrixSize
    maps = {
        "echotimes": (opts.TE * np.arange(1, opts.nTE + 1, dtype=np.float64)).astype(ft),
        "t2times": opts.T2_component_times().astype(ft),
    }
    if opts.SetFlipAngle is None:
        maps["refangleset"] = opts.flip_angles().astype(ft)
    maps["decaybasisset"] = decaybasisset(opts, device).astype(ft, order="F")
    out = {k: _nan(shape3, ft) for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha")}
    if alpha is not None:
        alpha = np.asfortranarray(alpha, dtype=ft)
        assert alpha.shape == shape3
        out["alpha"][...] = alpha
    if opts.SaveResidualNorm:
        out["resnorm"] = _nan(shape3, ft)
    if opts.SaveDecayCurve:
        out["decaycurve"] = _nan((*shape3, opts.nTE), ft)
    if opts.SaveRegParam:
        out["mu"] = _nan(shape3, ft)
        out["chi2factor"] = _nan(shape3, ft)
    if opts.SaveNNLSBasis:
        out["decaybasis"] = (_nan((*shape3, opts.nTE, opts.nT2), ft) if opts.SetFlipAngle is None
                             else np.empty((opts.nTE, opts.nT2), order="F", dtype=ft))
    dist = _nan((*shape3, opts.nT2), ft)
    oc = t2map_opts_c(opts)
    keys = ("gdn", "ggm", "gva", "fnr", "snr", "alpha", "resnorm", "decaycurve", "mu", "chi2factor", "decaybasis")
    if ft is np.float32:
        oo = _lib.T2mapOutF32C(_ptr(dist), *(_ptr(out.get(k)) for k in keys))
        rc = _lib.load().dmi_t2map_f32(ctx.handle, _ptr(image), C.byref(oc), _ptr(alpha), C.byref(oo))
    else:
        oo = _lib.T2mapOutC(_ptr(dist), *(_ptr(out.get(k)) for k in keys))
        rc = _lib.load().dmi_t2map_f64(ctx.handle, _ptr(image), C.byref(oc), _ptr(alpha), C.byref(oo))
    _lib.check(rc, ctx.handle)
    st = ctx.stats()
    if st["voxels_failed"]:
        warnings.warn(f"{st['voxels_failed']} voxels were left NaN: a per-voxel search cache overflowed.")
    if st["voxels"] == 0 and not opts.Silent:
        warnings.warn(f"No voxels found with first-echo signal intensity above threshold {opts.Threshold}.\n"
                      "Consider lowering the threshold or checking the input data.")
    maps.update(out)
    return maps, dist


def T2partSEcorr(dist: np.ndarray, opts: Optional[T2partOptions] = None, *, device: Optional[int] = None, **kwargs):
    """T2partSEcorr(dist, opts) (src/T2partSEcorr.jl:40-71) -> {"sfr", "sgm", "mfr", "mgm"}."""
    ft = _float_type(dist)
    dist = np.asfortranarray(dist, dtype=ft)
    if opts is None:
        opts = T2partOptions.from_dist(dist, **kwargs)
    elif kwargs:
        raise TypeError("pass either opts or keyword options, not both")
    assert dist.shape == (*opts.MatrixSize, opts.nT2)
    ctx = _lib.context(device)
    out = {k: _nan(opts.MatrixSize, ft) for k in ("sfr", "sgm", "mfr", "mgm")}
    oc = t2part_opts_c(opts)
    fn = _lib.load().dmi_t2part_f32 if ft is np.float32 else _lib.load().dmi_t2part_f64
    rc = fn(ctx.handle, _ptr(dist), C.byref(oc), _ptr(out["sfr"]), _ptr(out["sgm"]), _ptr(out["mfr"]), _ptr(out["mgm"]))
    _lib.check(rc, ctx.handle)
    return out


def synth_mse(shape, nTE: int, TE: float, *, T1: float = 1.0, ncomp: int = 3, mask_fraction: float = 0.6,
              seed: int = 0, nthreads: int = 0):
    """Seeded synthetic multi-echo spin-echo volume (X,Y,Z,nTE), Fortran order; returns (image, n_in_mask)."""
    img = np.empty((*shape, nTE), order="F")
    n = _lib.load().dmi_synth_mse(_ptr(img), int(shape[0]), int(shape[1]), int(shape[2]), int(nTE), float(TE),
                                  float(T1), int(ncomp), float(mask_fraction), int(seed) & (2**64 - 1), int(nthreads))
    if n < 0:
        raise ValueError("invalid synth_mse arguments")
    return img, int(n)
