"""ctypes binding to the FP64 CPU oracle (oracle/decaes_oracle.c).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIB_PATH = ORACLE_DIR / "_build" / "liborc.so"

REG = {"none": 0, "lcurve": 1, "gcv": 2, "chi2": 3, "mdp": 4}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
FUN1 = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
FUN2 = C.CFUNCTYPE(None, C.c_double, C.c_void_p, _dp)


class T2mapOpts(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("nTE", C.c_int32), ("nT2", C.c_int32),
        ("TE", C.c_double), ("T2Range", C.c_double * 2), ("T1", C.c_double), ("Threshold", C.c_double),
        ("MinRefAngle", C.c_double), ("nRefAngles", C.c_int32), ("nRefAnglesMin", C.c_int32),
        ("Reg", C.c_int32), ("legacy", C.c_int32),
        ("Chi2Factor", C.c_double), ("NoiseLevel", C.c_double), ("RefConAngle", C.c_double),
        ("SetFlipAngle", C.c_double),
    ]


class T2partOpts(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("nT2", C.c_int32), ("_pad0", C.c_int32),
        ("T2Range", C.c_double * 2), ("SPWin", C.c_double * 2), ("MPWin", C.c_double * 2), ("Sigmoid", C.c_double),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_nnls.restype = C.c_double
        L.orc_nnls_tikh.restype = C.c_double
        L.orc_sind.restype = C.c_double
        L.orc_sind.argtypes = [C.c_double]
        L.orc_cosd.restype = C.c_double
        L.orc_cosd.argtypes = [C.c_double]
        L.orc_erfinv.restype = C.c_double
        L.orc_erfinv.argtypes = [C.c_double]
        L.orc_gcv_dof.restype = C.c_double
        L.orc_lcurve_corner.restype = C.c_double
        L.orc_flip_search.restype = C.c_double
        L.orc_t2map.restype = C.c_int64
        L.orc_t2map_subset.restype = C.c_int64
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"] or a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def linrange(a, b, n):
    out = np.empty(n)
    lib().orc_linrange(C.c_double(a), C.c_double(b), C.c_int(n), _p(out))
    return out


def logrange(a, b, n):
    out = np.empty(n)
    lib().orc_logrange(C.c_double(a), C.c_double(b), C.c_int(n), _p(out))
    return out


def epg_const(ETL, alpha, TE, T2, T1):
    dc = np.empty(ETL)
    lib().orc_epg_const(_p(dc), C.c_int(ETL), C.c_double(alpha), C.c_double(TE), C.c_double(T2), C.c_double(T1))
    return dc


def epg_general(ETL, alpha, TE, T2, T1, beta):
    dc = np.empty(ETL)
    lib().orc_epg_general(_p(dc), C.c_int(ETL), C.c_double(alpha), C.c_double(TE), C.c_double(T2), C.c_double(T1),
                          C.c_double(beta))
    return dc


def epg_basic(ETL, alpha, TE, T2, T1, beta):
    dc = np.empty(ETL)
    lib().orc_epg_basic(_p(dc), C.c_int(ETL), C.c_double(alpha), C.c_double(TE), C.c_double(T2), C.c_double(T1),
                        C.c_double(beta))
    return dc


def epg_dalpha(ETL, alpha, TE, T2, T1, beta=180.0):
    dc, ddc = np.empty(ETL), np.empty(ETL)
    lib().orc_epg_dalpha(_p(dc), _p(ddc), C.c_int(ETL), C.c_double(alpha), C.c_double(TE), C.c_double(T2),
                         C.c_double(T1), C.c_double(beta))
    return dc, ddc


def nnls(A, b):
    A = _f(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = A.shape
    x = np.empty(n)
    w = np.empty(n)
    idx = np.empty(n, dtype=np.int32)
    nsetp, mode = C.c_int(), C.c_int()
    r = lib().orc_nnls(_p(A), _p(b), C.c_int(m), C.c_int(n), _p(x), C.byref(nsetp), C.byref(mode),
                       idx.ctypes.data_as(_ip), _p(w))
    return dict(x=x, rnorm=r, nsetp=nsetp.value, mode=mode.value, idx=idx, w=w)


def nnls_tikh(A, b, mu):
    A = _f(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = A.shape
    x = np.empty(n)
    nsetp, mode = C.c_int(), C.c_int()
    r = lib().orc_nnls_tikh(_p(A), _p(b), C.c_int(m), C.c_int(n), C.c_double(mu), _p(x), C.byref(nsetp),
                            C.byref(mode))
    return dict(x=x, rnorm=r, nsetp=nsetp.value, mode=mode.value)


def lsqnonneg(A, b, reg, param=None):
    A = _f(A)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m, n = A.shape
    x = np.empty(n)
    mu, chi2 = C.c_double(), C.c_double()
    L = lib()
    args = [_p(A), _p(b), C.c_int(m), C.c_int(n)]
    if reg == "none":
        L.orc_lsqnonneg_none(*args, _p(x), C.byref(mu), C.byref(chi2))
    elif reg == "chi2":
        L.orc_lsqnonneg_chi2(*args, C.c_double(param), _p(x), C.byref(mu), C.byref(chi2))
    elif reg == "chi2_legacy":
        L.orc_lsqnonneg_chi2_legacy(*args, C.c_double(param), _p(x), C.byref(mu), C.byref(chi2))
    elif reg == "mdp":
        L.orc_lsqnonneg_mdp(*args, C.c_double(param), _p(x), C.byref(mu), C.byref(chi2))
    elif reg == "gcv":
        L.orc_lsqnonneg_gcv(*args, _p(x), C.byref(mu), C.byref(chi2))
    elif reg == "lcurve":
        L.orc_lsqnonneg_lcurve(*args, _p(x), C.byref(mu), C.byref(chi2))
    else:
        raise ValueError(reg)
    return x, mu.value, chi2.value


def svdvals(A):
    A = _f(A)
    m, n = A.shape
    s = np.empty(min(m, n))
    lib().orc_svdvals(_p(A), C.c_int(m), C.c_int(n), _p(s))
    return s


def gcv_dof(m, n, gamma, lam):
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    return lib().orc_gcv_dof(C.c_int(m), C.c_int(n), _p(g), C.c_int(len(g)), C.c_double(lam))


def brent_root(f, x0, x1, xatol=0.0, xrtol=0.0, ftol=0.0, maxiters=100):
    cb = FUN1(lambda x, _ctx: float(f(x)))
    xo, fo = C.c_double(), C.c_double()
    lib().orc_brent_root(cb, None, C.c_double(x0), C.c_double(x1), C.c_double(xatol), C.c_double(xrtol),
                         C.c_double(ftol), C.c_int(maxiters), C.byref(xo), C.byref(fo))
    return xo.value, fo.value


def bracket_root_monotonic(f, a, delta, dilate=1.0, mono=1.0, maxiters=100):
    cb = FUN1(lambda x, _ctx: float(f(x)))
    out = (C.c_double * 4)()
    lib().orc_bracket_root_monotonic(cb, None, C.c_double(a), C.c_double(delta), C.c_double(dilate),
                                     C.c_double(mono), C.c_int(maxiters), out)
    return tuple(out)


def brent_minimize(f, x1, x2, xrtol=1.4901161193847656e-8, xatol=1.4901161193847656e-8, maxiters=100):
    cb = FUN1(lambda x, _ctx: float(f(x)))
    xo, fo = C.c_double(), C.c_double()
    lib().orc_brent_minimize(cb, None, C.c_double(x1), C.c_double(x2), C.c_double(xrtol), C.c_double(xatol),
                             C.c_int(maxiters), C.byref(xo), C.byref(fo))
    return xo.value, fo.value


def lcurve_corner(f, xlow, xhigh, xtol=1e-4, Ptol=1e-4, Ctol=1e-4, backtracking=True):
    def _cb(x, _ctx, out):
        a, b = f(x)
        out[0], out[1] = a, b

    cb = FUN2(_cb)
    L = lib()
    L.orc_lcurve_corner_bt.restype = C.c_double
    return L.orc_lcurve_corner_bt(cb, None, C.c_double(xlow), C.c_double(xhigh), C.c_double(xtol),
                                  C.c_double(Ptol), C.c_double(Ctol), C.c_int(1 if backtracking else 0))


def hermite_minimize(a, b, u0, u1, m0, m1):
    out = (C.c_double * 2)()
    lib().orc_hermite_minimize(*(C.c_double(v) for v in (a, b, u0, u1, m0, m1)), out)
    return out[0], out[1]


def flip_search(As, dAs, alphas, b, mineval):
    As = np.asfortranarray(As, dtype=np.float64)
    dAs = np.asfortranarray(dAs, dtype=np.float64)
    m, n, K = As.shape
    ne = C.c_int()
    a = lib().orc_flip_search(_p(As), _p(dAs), _p(np.ascontiguousarray(alphas, dtype=np.float64)), C.c_int(m),
                              C.c_int(n), C.c_int(K), C.c_int(mineval),
                              _p(np.ascontiguousarray(b, dtype=np.float64)), C.byref(ne))
    return a, ne.value


LOSS = C.CFUNCTYPE(None, C.c_void_p, C.c_int, _dp, _dp)


def surrogate_search(alphas, mineval, loss, legacy=False):
    """The flip-angle searcher on loss(alpha) -> (u, du/dalpha) over the grid alphas (with legacy, the cubic spline
    surrogate sampled on a 0.001 grid; du unused). Returns (x, trace), trace the 1-based grid indices in evaluation
    order."""
    alphas = np.ascontiguousarray(alphas, dtype=np.float64)
    K = len(alphas)

    def _cb(_ctx, I, u, du):
        uu, dd = loss(float(alphas[I - 1]))
        u[0], du[0] = uu, dd

    cb = LOSS(_cb)
    trace = np.zeros(K, dtype=np.int32)
    nev = C.c_int(0)
    L = lib()
    fn = L.orc_surrogate_search_legacy if legacy else L.orc_surrogate_search
    fn.restype = C.c_double
    x = fn(_p(alphas), C.c_int(K), C.c_int(mineval), cb, None, trace.ctypes.data_as(_ip), C.byref(nev))
    return x, [int(i) for i in trace[:nev.value]]


def spline_interp(x, y, k):
    """Dierckx.Spline1D(x, y; k, s = 0): (knots, coefficients)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = len(x)
    t = np.empty(m + k + 1)
    c = np.empty(m)
    n = lib().orc_spline_interp(_p(x), _p(y), C.c_int(m), C.c_int(k), _p(t), _p(c))
    assert n == m + k + 1, n
    return t, c


def spline_eval(t, c, k, x):
    L = lib()
    L.orc_spline_eval.restype = C.c_double
    t = np.ascontiguousarray(t, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    return np.array([L.orc_spline_eval(_p(t), C.c_int(len(t)), _p(c), C.c_int(k), C.c_double(float(v)))
                     for v in np.atleast_1d(x)])


def jlrange(start, step, stop, cap=None):
    """Julia's Float64 range start:step:stop as an array (its first cap elements when cap is given)."""
    L = lib()
    L.orc_jlrange.restype = C.c_int64
    n = L.orc_jlrange(C.c_double(start), C.c_double(step), C.c_double(stop), None, C.c_int64(0))
    cap = n if cap is None else min(n, cap)
    out = np.empty(max(cap, 1))
    L.orc_jlrange(C.c_double(start), C.c_double(step), C.c_double(stop), _p(out), C.c_int64(cap))
    return out[:cap], n


def spline_opt_legacy(x, y):
    out = (C.c_double * 2)()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_spline_opt_legacy(_p(x), _p(y), C.c_int(len(x)), out)
    return out[0], out[1]


def spline_root_legacy(x, y, value):
    L = lib()
    L.orc_spline_root_legacy.restype = C.c_double
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return L.orc_spline_root_legacy(_p(x), _p(y), C.c_int(len(x)), C.c_double(value))


def minimal_bounding_box(seen_indices, alphas, x):
    alphas = np.ascontiguousarray(alphas, dtype=np.float64)
    K = len(alphas)
    seen = np.zeros(K + 1, dtype=np.uint8)
    seen[list(seen_indices)] = 1
    lo, hi = C.c_int(0), C.c_int(0)
    lib().orc_minimal_bounding_box(seen.ctypes.data_as(C.POINTER(C.c_ubyte)), _p(alphas), C.c_int(K), C.c_double(x),
                                   C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def sorted_corners(alphas, lo, hi, x):
    alphas = np.ascontiguousarray(alphas, dtype=np.float64)
    out = (C.c_int * 2)()
    lib().orc_sorted_corners(_p(alphas), C.c_int(lo), C.c_int(hi), C.c_double(x), out)
    return out[0], out[1]


def make_t2map_opts(shape, nTE, TE, nT2, T2Range, Reg, T1=1.0, Threshold=0.0, MinRefAngle=50.0, nRefAngles=64,
                    nRefAnglesMin=None, Chi2Factor=None, NoiseLevel=None, RefConAngle=180.0, SetFlipAngle=None,
                    legacy=False):
    o = T2mapOpts()
    o.nx, o.ny, o.nz = shape
    o.nTE, o.nT2, o.TE = nTE, nT2, TE
    o.T2Range[0], o.T2Range[1] = T2Range
    o.T1, o.Threshold, o.MinRefAngle = T1, Threshold, MinRefAngle
    o.nRefAngles = nRefAngles
    o.nRefAnglesMin = min(5, nRefAngles) if nRefAnglesMin is None else nRefAnglesMin
    o.Reg = REG[Reg]
    o.Chi2Factor = math.nan if Chi2Factor is None else Chi2Factor
    o.NoiseLevel = math.nan if NoiseLevel is None else NoiseLevel
    o.RefConAngle = RefConAngle
    o.SetFlipAngle = math.nan if SetFlipAngle is None else SetFlipAngle
    o.legacy = int(bool(legacy))
    return o


def t2map(image, opts: T2mapOpts, alpha_in=None, nthreads=None, save_all=True):
    """Run the oracle pipeline. Returns (maps dict, dist) with NaN where voxels were skipped."""
    image = np.asfortranarray(image, dtype=np.float64)
    shape = image.shape[:3]
    m, n = opts.nTE, opts.nT2
    K = opts.nRefAngles if math.isnan(opts.SetFlipAngle) else 1
    nan3 = lambda: np.full(shape, np.nan, order="F")
    maps = {k: nan3() for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha")}
    dist = np.full(shape + (n,), np.nan, order="F")
    extra = {}
    if save_all:
        extra = dict(resnorm=nan3(), mu=nan3(), chi2factor=nan3(),
                     decaycurve=np.full(shape + (m,), np.nan, order="F"))
    basisset = np.empty((m, n, K) if K > 1 else (m, n), order="F")
    if alpha_in is not None:
        alpha_in = np.asfortranarray(alpha_in, dtype=np.float64)
        maps["alpha"][...] = alpha_in
    nthreads = nthreads or os.cpu_count() or 1
    nv = lib().orc_t2map(C.byref(opts), _p(image), _p(alpha_in), _p(dist), _p(maps["gdn"]), _p(maps["ggm"]),
                         _p(maps["gva"]), _p(maps["fnr"]), _p(maps["snr"]), _p(maps["alpha"]),
                         _p(extra.get("resnorm")), _p(extra.get("decaycurve")), _p(extra.get("mu")),
                         _p(extra.get("chi2factor")), None, _p(basisset), C.c_int(nthreads))
    if nv < 0:
        raise RuntimeError("oracle t2map failed")
    maps.update(extra)
    maps["decaybasisset"] = basisset
    return maps, dist, nv


def t2map_subset(image, opts: T2mapOpts, vox, nthreads=1):
    image = np.asfortranarray(image, dtype=np.float64)
    shape = image.shape[:3]
    n = opts.nT2
    vox = np.ascontiguousarray(vox, dtype=np.int64)
    nan3 = lambda: np.full(shape, np.nan, order="F")
    maps = {k: nan3() for k in ("gdn", "ggm", "gva", "fnr", "snr", "alpha", "mu", "chi2factor")}
    dist = np.full(shape + (n,), np.nan, order="F")
    lib().orc_t2map_subset(C.byref(opts), _p(image), vox.ctypes.data_as(_i64p), C.c_int64(len(vox)), _p(dist),
                           _p(maps["gdn"]), _p(maps["ggm"]), _p(maps["gva"]), _p(maps["fnr"]), _p(maps["snr"]),
                           _p(maps["alpha"]), _p(maps["mu"]), _p(maps["chi2factor"]), C.c_int(nthreads))
    return maps, dist


def t2part(dist, T2Range, SPWin, MPWin, Sigmoid=None):
    dist = np.asfortranarray(dist, dtype=np.float64)
    o = T2partOpts()
    o.nx, o.ny, o.nz = dist.shape[:3]
    o.nT2 = dist.shape[3]
    o.T2Range[0], o.T2Range[1] = T2Range
    o.SPWin[0], o.SPWin[1] = SPWin
    o.MPWin[0], o.MPWin[1] = MPWin
    o.Sigmoid = math.nan if Sigmoid is None else Sigmoid
    shape = dist.shape[:3]
    out = {k: np.full(shape, np.nan, order="F") for k in ("sfr", "sgm", "mfr", "mgm")}
    rc = lib().orc_t2part(C.byref(o), _p(dist), _p(out["sfr"]), _p(out["sgm"]), _p(out["mfr"]), _p(out["mgm"]))
    if rc != 0:
        raise RuntimeError("oracle t2part failed")
    return out
