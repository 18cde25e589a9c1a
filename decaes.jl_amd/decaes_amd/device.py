"""Device-resident entry points: inputs and outputs already in HBM (torch tensors used as plain buffers).

A Julia column-major (X,Y,Z,k) array maps to a contiguous torch tensor of shape (k,Z,Y,X): same bytes.
torch only provides device memory and the current device here; all voxel work is in libdecaes_amd.so.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

from . import _lib
from .api import t2map_opts_c, t2part_opts_c
from .options import T2mapOptions, T2partOptions

_OPTIONAL = ("resnorm", "decaycurve", "mu", "chi2factor", "decaybasis")


def alloc_t2map_outputs(opts: T2mapOptions, device="cuda"):
    """NaN-filled output tensors in column-major-equivalent shapes."""
    import torch

    X, Y, Z = opts.MatrixSize
    f = lambda *s: torch.full(s, float("nan"), dtype=torch.float64, device=device)
    out = {k: f(Z, Y, X) for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha")}
    out["dist"] = f(opts.nT2, Z, Y, X)
    if opts.SaveResidualNorm:
        out["resnorm"] = f(Z, Y, X)
    if opts.SaveDecayCurve:
        out["decaycurve"] = f(opts.nTE, Z, Y, X)
    if opts.SaveRegParam:
        out["mu"] = f(Z, Y, X)
        out["chi2factor"] = f(Z, Y, X)
    if opts.SaveNNLSBasis:
        out["decaybasis"] = f(opts.nT2, opts.nTE, Z, Y, X) if opts.SetFlipAngle is None else f(opts.nT2, opts.nTE)
    return out


def _ready(t):
    """The library runs on its own HIP stream: work torch queued on its current stream for the arrays (an input
    computed just before the call, outputs just filled) must be finished before the library reads them."""
    import torch

    torch.cuda.current_stream(t.device).synchronize()


def _dptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous() and str(t.dtype) == "torch.float64"
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def t2map_device(image, opts: T2mapOptions, out: dict, alpha=None, ctx: Optional[_lib.Context] = None,
                 noalloc: bool = False) -> dict:
    """dmi_t2map_f64_device on torch tensors (dmi_t2map_f64_device_noalloc with noalloc=True: no device allocation
    in the call); returns the call's device counters (dmi_last_stats)."""
    ctx = ctx or _lib.context(image.device.index or 0)
    _ready(image)
    oc = t2map_opts_c(opts)
    oo = _lib.T2mapOutC(_dptr(out["dist"]), *(_dptr(out.get(k)) for k in (
        "gdn", "ggm", "gva", "fnr", "snr", "alpha", *_OPTIONAL)))
    L = _lib.load()
    fn = L.dmi_t2map_f64_device_noalloc if noalloc else L.dmi_t2map_f64_device
    rc = fn(ctx.handle, C.c_void_p(image.data_ptr()), C.byref(oc),
                                          C.c_void_p(alpha.data_ptr()) if alpha is not None else None, C.byref(oo))
    _lib.check(rc, ctx.handle)
    return ctx.stats()


def t2part_device(dist, opts: T2partOptions, out: dict, ctx: Optional[_lib.Context] = None) -> None:
    ctx = ctx or _lib.context(dist.device.index or 0)
    _ready(dist)
    oc = t2part_opts_c(opts)
    v = C.c_void_p
    rc = _lib.load().dmi_t2part_f64_device(ctx.handle, v(dist.data_ptr()), C.byref(oc), v(out["sfr"].data_ptr()),
                                           v(out["sgm"].data_ptr()), v(out["mfr"].data_ptr()), v(out["mgm"].data_ptr()))
    _lib.check(rc, ctx.handle)
