"""Request loop on one GPU (dmi_server_* in include/decaes_amd.h).

The counterpart of the reference's warm server (api/decaes_server.jl:95-131), which keeps a Julia process alive and
runs one CLI call at a time. Here requests are T2mapSEcorr calls (optionally followed by T2partSEcorr on the
distribution) that stay on the device while they run: they are admitted while their device footprint fits an HBM
budget, split into voxel chunks and interleaved on the GPU with the other requests in flight, so a request that
arrives while others run starts at the next chunk boundary. Results equal those of T2mapSEcorr on the same input.

    srv = Server()
    h = srv.submit(image, T2mapOptions(...), T2partOptions(...))
    maps, dist, parts, info = srv.wait(h)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .api import _nan, _ptr, t2map_opts_c, t2part_opts_c
from .options import T2mapOptions, T2partOptions


@dataclass
class Request:
    id: int
    opts: T2mapOptions
    image: np.ndarray
    out: dict
    dist: np.ndarray
    parts: Optional[dict]
    keep: list = field(default_factory=list)  # ctypes structs and arrays the library reads until wait()
    done: bool = False


def request_arrays(image: np.ndarray, opts: T2mapOptions, alpha: Optional[np.ndarray] = None):
    """The host arrays of one request -> (image, alpha, maps, dist): the image and B1 map as column-major FP64 of the
    options' sizes (the library reads exactly that many doubles from each), and the NaN-filled outputs, with the B1
    map in maps["alpha"] (load_B1map!, src/T2mapSEcorr.jl:56-61)."""
    image = np.asfortranarray(image, dtype=np.float64)
    assert image.shape == (*opts.MatrixSize, opts.nTE), f"image shape {image.shape} != {(*opts.MatrixSize, opts.nTE)}"
    shape3 = opts.MatrixSize
    out = {k: _nan(shape3) for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha")}
    if alpha is not None:
        alpha = np.asfortranarray(alpha, dtype=np.float64)
        assert alpha.shape == tuple(shape3), f"B1 map shape {alpha.shape} != {tuple(shape3)}"
        out["alpha"][...] = alpha
    if opts.SaveResidualNorm:
        out["resnorm"] = _nan(shape3)
    if opts.SaveDecayCurve:
        out["decaycurve"] = _nan((*shape3, opts.nTE))
    if opts.SaveRegParam:
        out["mu"] = _nan(shape3)
        out["chi2factor"] = _nan(shape3)
    if opts.SaveNNLSBasis:
        out["decaybasis"] = (_nan((*shape3, opts.nTE, opts.nT2)) if opts.SetFlipAngle is None
                             else np.empty((opts.nTE, opts.nT2), order="F"))
    return image, alpha, out, _nan((*shape3, opts.nT2))


class Server:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, max_streams: int = 0, chunk_voxels: int = 0,
                 order: str = "shortest"):
        """order: "shortest" (the request with the fewest voxels left gets a stream's next chunk) or "turns" (a
        stream's requests take turns, one chunk each); dmi_server_config.order."""
        if order not in ("shortest", "turns"):
            raise ValueError(f"order must be 'shortest' or 'turns', not {order!r}")
        L = _lib.load()
        cfg = _lib.ServerConfigC(int(hbm_budget_bytes), int(max_streams), int(chunk_voxels),
                                 0 if order == "shortest" else 1, 0)
        h = C.c_void_p()
        rc = L.dmi_server_create(C.byref(h), int(device), C.byref(cfg))
        if rc != _lib.DMI_OK:
            raise _lib.DmiError(rc, "dmi_server_create failed (no gfx950 device?)")
        self.handle = h
        # every request not yet returned by wait(): the library reads and writes its host arrays until then, so
        # they are kept alive here whether or not the caller keeps the Request
        self._pending: dict[int, Request] = {}

    def _error(self) -> str:
        return _lib.load().dmi_server_last_error(self.handle).decode(errors="replace")

    def submit(self, image: np.ndarray, opts: T2mapOptions, part_opts: Optional[T2partOptions] = None,
               alpha: Optional[np.ndarray] = None) -> Request:
        image, alpha, out, dist = request_arrays(image, opts, alpha)
        shape3 = opts.MatrixSize
        oc = t2map_opts_c(opts)
        oo = _lib.T2mapOutC(_ptr(dist), *(_ptr(out.get(k)) for k in (
            "gdn", "ggm", "gva", "fnr", "snr", "alpha", "resnorm", "decaycurve", "mu", "chi2factor", "decaybasis")))
        parts, pc = None, None
        if part_opts is not None:
            parts = {k: _nan(shape3) for k in ("sfr", "sgm", "mfr", "mgm")}
            pc = t2part_opts_c(part_opts)
        rid = C.c_int64(0)
        rc = _lib.load().dmi_server_submit(
            self.handle, _ptr(image), C.byref(oc), _ptr(alpha), C.byref(oo), C.byref(pc) if pc is not None else None,
            *((_ptr(parts[k]) for k in ("sfr", "sgm", "mfr", "mgm")) if parts else (None,) * 4), C.byref(rid))
        if rc == _lib.DMI_EINVAL:
            raise AssertionError(self._error())
        if rc != _lib.DMI_OK:
            raise _lib.DmiError(rc, self._error())
        req = Request(rid.value, opts, image, out, dist, parts, keep=[oc, oo, pc, alpha])
        self._pending[req.id] = req
        return req

    def wait(self, req: Request):
        """Blocks until the request finished -> (maps, dist, parts or None, info dict). Raises on failure."""
        info = _lib.RequestInfoC()
        rc = _lib.load().dmi_server_wait(self.handle, req.id, C.byref(info))
        req.done = True
        self._pending.pop(req.id, None)
        inf = {k: getattr(info, k) for k, _ in _lib.RequestInfoC._fields_ if not k.startswith("_")}
        if rc != _lib.DMI_OK:
            msg = _lib.load().dmi_server_request_error(self.handle, req.id).decode(errors="replace")
            raise _lib.DmiError(rc, msg)
        return req.out, req.dist, req.parts, inf

    def poll(self, req: Request) -> bool:
        d = C.c_int32(0)
        _lib.load().dmi_server_poll(self.handle, req.id, C.byref(d))
        return bool(d.value)

    def stats(self) -> dict:
        s = _lib.ServerStatsC()
        _lib.load().dmi_server_get_stats(self.handle, C.byref(s))
        return {k: getattr(s, k) for k, _ in _lib.ServerStatsC._fields_}

    def close(self):
        if self.handle:
            _lib.load().dmi_server_destroy(self.handle)  # waits for every submitted request
            self.handle = None
        self._pending = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
