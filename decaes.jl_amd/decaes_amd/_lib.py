"""ctypes binding of libdecaes_amd.so (include/decaes_amd.h).

The shared library is the product: every T2mapSEcorr / T2partSEcorr voxel runs in its gfx950 kernels.
Loading fails loudly when the library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdecaes_amd.so"

DMI_OK = 0
DMI_EINVAL = -1
DMI_EUNSUPPORTED = -2
DMI_ENOMEM = -3
DMI_EHIP = -4
DMI_ENODEV = -5
DMI_ECAPACITY = -6

REG = {"none": 0, "lcurve": 1, "gcv": 2, "chi2": 3, "mdp": 4}
PREC = {"fp64": 0, "mixed": 1}
SCHED = {"auto": 0, "lds": 1, "split": 2}

_dp = C.POINTER(C.c_double)

# Every symbol include/decaes_amd.h declares.
EXPORTS = (
    "dmi_create", "dmi_destroy", "dmi_set_stream", "dmi_last_error", "dmi_t2map_f64", "dmi_t2map_f64_device",
    "dmi_epg_basis_set",
    "dmi_t2part_f64", "dmi_t2part_f64_device", "dmi_last_stats", "dmi_synth_mse",
    "dmi_t2map_f32", "dmi_t2part_f32", "dmi_t2map_f64_device_noalloc",
    "dmi_server_create", "dmi_server_destroy", "dmi_server_submit", "dmi_server_wait", "dmi_server_poll",
    "dmi_server_get_stats", "dmi_server_last_error", "dmi_server_request_error",
)


class T2mapOptsC(C.Structure):
    _fields_ = [
        ("MatrixSize", C.c_int64 * 3), ("nTE", C.c_int32), ("nT2", C.c_int32), ("TE", C.c_double),
        ("T2Range", C.c_double * 2), ("T1", C.c_double), ("Threshold", C.c_double), ("MinRefAngle", C.c_double),
        ("nRefAngles", C.c_int32), ("nRefAnglesMin", C.c_int32), ("Reg", C.c_int32), ("legacy", C.c_int32),
        ("Chi2Factor", C.c_double), ("NoiseLevel", C.c_double), ("RefConAngle", C.c_double),
        ("SetFlipAngle", C.c_double), ("SaveResidualNorm", C.c_int32), ("SaveDecayCurve", C.c_int32),
        ("SaveRegParam", C.c_int32), ("SaveNNLSBasis", C.c_int32), ("precision", C.c_int32), ("schedule", C.c_int32),
    ]


class T2mapOutC(C.Structure):
    _fields_ = [(k, _dp) for k in ("dist", "gdn", "ggm", "gva", "fnr", "snr", "alpha", "resnorm", "decaycurve", "mu",
                                   "chi2factor", "decaybasis")]


_fp = C.POINTER(C.c_float)


class T2mapOutF32C(C.Structure):
    _fields_ = [(k, _fp) for k in ("dist", "gdn", "ggm", "gva", "fnr", "snr", "alpha", "resnorm", "decaycurve", "mu",
                                   "chi2factor", "decaybasis")]


class T2partOptsC(C.Structure):
    _fields_ = [
        ("MatrixSize", C.c_int64 * 3), ("nT2", C.c_int32), ("legacy", C.c_int32), ("T2Range", C.c_double * 2),
        ("SPWin", C.c_double * 2), ("MPWin", C.c_double * 2), ("Sigmoid", C.c_double),
    ]


class StatsC(C.Structure):
    _fields_ = [("voxels", C.c_int64), ("nnls_solves", C.c_int64), ("fp64_flops", C.c_int64),
                ("fp32_flops", C.c_int64), ("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("voxels_failed", C.c_int64), ("lds_rows", C.c_int32), ("workgroups_per_cu", C.c_int32)]


class ServerConfigC(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_streams", C.c_int32), ("chunk_voxels", C.c_int32),
                ("order", C.c_int32), ("reserved", C.c_int32)]


class RequestInfoC(C.Structure):
    _fields_ = [("status", C.c_int32), ("workgroups_per_cu", C.c_int32), ("voxels", C.c_int64),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double), ("trial_voxels", C.c_int64)]


class ServerStatsC(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("voxels", C.c_int64),
                ("busy_ms", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double)]


_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load libdecaes_amd.so; raises ImportError if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(or decaes.jl_amd/decaes_amd/_build.py). There is no CPU fallback.")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.dmi_create.argtypes = [C.POINTER(vp), C.c_int, C.c_uint32]
        L.dmi_destroy.argtypes = [vp]
        L.dmi_set_stream.argtypes = [vp, vp]
        L.dmi_last_error.argtypes = [vp]
        L.dmi_last_error.restype = C.c_char_p
        L.dmi_t2map_f64.argtypes = [vp, _dp, C.POINTER(T2mapOptsC), _dp, C.POINTER(T2mapOutC)]
        L.dmi_t2map_f64_device.argtypes = [vp, vp, C.POINTER(T2mapOptsC), vp, C.POINTER(T2mapOutC)]
        L.dmi_t2map_f64_device_noalloc.argtypes = [vp, vp, C.POINTER(T2mapOptsC), vp, C.POINTER(T2mapOutC)]
        L.dmi_t2map_f32.argtypes = [vp, _fp, C.POINTER(T2mapOptsC), _fp, C.POINTER(T2mapOutF32C)]
        L.dmi_t2part_f32.argtypes = [vp, _fp, C.POINTER(T2partOptsC), _fp, _fp, _fp, _fp]
        L.dmi_epg_basis_set.argtypes = [vp, C.POINTER(T2mapOptsC), _dp]
        L.dmi_t2part_f64.argtypes = [vp, _dp, C.POINTER(T2partOptsC), _dp, _dp, _dp, _dp]
        L.dmi_t2part_f64_device.argtypes = [vp, vp, C.POINTER(T2partOptsC), vp, vp, vp, vp]
        L.dmi_last_stats.argtypes = [vp, C.POINTER(StatsC)]
        L.dmi_synth_mse.argtypes = [_dp, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_double, C.c_double,
                                    C.c_int32, C.c_double, C.c_uint64, C.c_int32]
        L.dmi_synth_mse.restype = C.c_int64
        L.dmi_server_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(ServerConfigC)]
        L.dmi_server_destroy.argtypes = [vp]
        L.dmi_server_submit.argtypes = [vp, _dp, C.POINTER(T2mapOptsC), _dp, C.POINTER(T2mapOutC),
                                        C.POINTER(T2partOptsC), _dp, _dp, _dp, _dp, C.POINTER(C.c_int64)]
        L.dmi_server_wait.argtypes = [vp, C.c_int64, C.POINTER(RequestInfoC)]
        L.dmi_server_poll.argtypes = [vp, C.c_int64, C.POINTER(C.c_int32)]
        L.dmi_server_get_stats.argtypes = [vp, C.POINTER(ServerStatsC)]
        L.dmi_server_last_error.argtypes = [vp]
        L.dmi_server_last_error.restype = C.c_char_p
        L.dmi_server_request_error.argtypes = [vp, C.c_int64]
        L.dmi_server_request_error.restype = C.c_char_p
        _lib = L
        return L


class DmiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(code: int, ctx=None):
    if code == DMI_OK:
        return
    msg = load().dmi_last_error(ctx).decode(errors="replace")
    if code == DMI_EINVAL:
        raise AssertionError(msg)
    raise DmiError(code, msg)


class Context:
    """One HIP stream and its device buffers on one GPU (dmi_create / dmi_destroy)."""

    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        check(L.dmi_create(C.byref(h), int(device), 0), None)
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            load().dmi_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream=None):
        """Runs later calls on `stream` (a torch.cuda.Stream or a raw hipStream_t handle), or on the context's own
        stream when None (dmi_set_stream)."""
        h = None if stream is None else getattr(stream, "cuda_stream", stream)
        check(load().dmi_set_stream(self.handle, C.c_void_p(h) if h else None), self.handle)

    def stats(self) -> dict:
        s = StatsC()
        check(load().dmi_last_stats(self.handle, C.byref(s)), self.handle)
        return {k: getattr(s, k) for k, _ in StatsC._fields_}


_contexts: dict[int, Context] = {}


def context(device: int | None = None) -> Context:
    if device is None:
        device = int(os.environ.get("DECAES_AMD_DEVICE", "0"))
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _contexts[device] = ctx
    return ctx
