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


def _check(items, device: int):
    """items: (name, tensor, numel). The library reads or writes exactly `numel` doubles at each tensor's address on
    `device`: anything else is refused here, before a kernel could touch memory outside a tensor. Sizes are checked
    for every tensor before devices."""
    for name, t, numel in items:
        if str(t.dtype) != "torch.float64" or not t.is_contiguous() or t.numel() != numel:
            raise ValueError(f"{name} must be a contiguous float64 tensor of {numel} elements, got {t.dtype} "
                             f"{tuple(t.shape)}{'' if t.is_contiguous() else ' (not contiguous)'}")
    for name, t, _ in items:
        if not t.is_cuda or (t.device.index or 0) != device:
            raise ValueError(f"{name} must be on cuda:{device}, got {t.device}")


def _check_t2map(image, opts: T2mapOptions, out: dict, alpha, device: int):
    X, Y, Z = opts.MatrixSize
    V = X * Y * Z
    if "dist" not in out:
        raise ValueError('out["dist"] is required')
    items = [("image", image, V * opts.nTE)]
    if alpha is not None:
        items.append(("alpha", alpha, V))
    sizes = dict(dist=V * opts.nT2, decaycurve=V * opts.nTE,
                 decaybasis=opts.nTE * opts.nT2 * (V if opts.SetFlipAngle is None else 1))
    for k, t in out.items():
        if t is not None and k in ("dist", "gdn", "ggm", "gva", "fnr", "snr", "alpha", *_OPTIONAL):
            items.append((f'out["{k}"]', t, sizes.get(k, V)))
    _check(items, device)


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


def _call(ctx, stream, fn, *args):
    """fn(ctx.handle, *args) on `stream` (dmi_set_stream) when one is given, else on the context's own stream."""
    if stream is not None:
        ctx.set_stream(stream)
    try:
        rc = fn(ctx.handle, *args)
    finally:
        if stream is not None:
            ctx.set_stream(None)
    _lib.check(rc, ctx.handle)


def t2map_device(image, opts: T2mapOptions, out: dict, alpha=None, ctx: Optional[_lib.Context] = None,
                 noalloc: bool = False, stream=None) -> dict:
    """dmi_t2map_f64_device on torch tensors (dmi_t2map_f64_device_noalloc with noalloc=True: no device allocation
    in the call); returns the call's device counters (dmi_last_stats). With `stream` (a torch.cuda.Stream) the call
    runs on that stream, after the work queued on it; while that stream is captured into a graph, a noalloc call
    records its launches there."""
    dev = ctx.device if ctx is not None else (image.device.index or 0)
    _check_t2map(image, opts, out, alpha, dev)
    ctx = ctx or _lib.context(dev)
    if stream is None:
        _ready(image)
    oc = t2map_opts_c(opts)
    oo = _lib.T2mapOutC(_dptr(out["dist"]), *(_dptr(out.get(k)) for k in (
        "gdn", "ggm", "gva", "fnr", "snr", "alpha", *_OPTIONAL)))
    L = _lib.load()
    fn = L.dmi_t2map_f64_device_noalloc if noalloc else L.dmi_t2map_f64_device
    _call(ctx, stream, fn, C.c_void_p(image.data_ptr()), C.byref(oc),
          C.c_void_p(alpha.data_ptr()) if alpha is not None else None, C.byref(oo))
    return ctx.stats()


def t2part_device(dist, opts: T2partOptions, out: dict, ctx: Optional[_lib.Context] = None, stream=None) -> None:
    """dmi_t2part_f64_device on torch tensors; `stream` as for t2map_device."""
    dev = ctx.device if ctx is not None else (dist.device.index or 0)
    X, Y, Z = opts.MatrixSize
    parts = [(f'out["{k}"]', out[k], X * Y * Z) for k in ("sfr", "sgm", "mfr", "mgm")]
    _check([("dist", dist, X * Y * Z * opts.nT2)] + parts, dev)
    ctx = ctx or _lib.context(dev)
    if stream is None:
        _ready(dist)
    oc = t2part_opts_c(opts)
    v = C.c_void_p
    _call(ctx, stream, _lib.load().dmi_t2part_f64_device, v(dist.data_ptr()), C.byref(oc), v(out["sfr"].data_ptr()),
          v(out["sgm"].data_ptr()), v(out["mfr"].data_ptr()), v(out["mgm"].data_ptr()))


def capture(image, opts: T2mapOptions, out: dict, alpha=None, t2part=None, ctx: Optional[_lib.Context] = None):
    """T2mapSEcorr (dmi_t2map_f64_device_noalloc) and, with t2part = (T2partOptions, outputs), T2partSEcorr on its
    distribution, recorded into a torch.cuda.CUDAGraph. Each graph.replay() fills the outputs with NaN (and
    out["alpha"] with the B1 map `alpha`, when given, as T2mapSEcorr's load_B1map! does, src/T2mapSEcorr.jl:56-61)
    and recomputes them from the current contents of `image` (and `alpha`), as fresh calls would.

    The calls run once outside capture first: that puts the grids of `opts` in the device's static scratch and the
    T2part tables on the context, which the graph reads at every replay. A later call with other options on the same
    device replaces them; call with the graph's options again before replaying."""
    import torch

    ctx = ctx or _lib.context(image.device.index or 0)
    t2map_device(image, opts, out, alpha, ctx, noalloc=True)
    if t2part is not None:
        t2part_device(out["dist"], t2part[0], t2part[1], ctx)
    torch.cuda.synchronize(image.device)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        s = torch.cuda.current_stream(image.device)
        for k, t in list(out.items()) + (list(t2part[1].items()) if t2part is not None else []):
            if alpha is not None and k == "alpha" and t is out.get("alpha"):
                t.copy_(alpha)
            else:
                t.fill_(float("nan"))
        t2map_device(image, opts, out, alpha, ctx, noalloc=True, stream=s)
        if t2part is not None:
            t2part_device(out["dist"], t2part[0], t2part[1], ctx, stream=s)
    return g
