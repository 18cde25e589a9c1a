"""MAT v7.3 files (HDF5) as MAT.jl's matwrite / matread handle them (src/main.jl:377-411 writes the results with
MAT.matwrite; load_image, src/main.jl:577-589, reads any MAT file with MAT.matread).

Layout written (the one MAT.jl writes for Float64 arrays, which is MATLAB's own v7.3 layout):
* an HDF5 file with a 512-byte user block holding the MAT header: the text "MATLAB 7.3 MAT-file", zero padding, and
  at byte offsets 125-127 the version 0x02 and the endian mark "IM";
* one dataset per variable at the root, IEEE float64, contiguous, with the dimensions reversed (HDF5 lists them
  slowest first, so the column-major bytes are stored unchanged), carrying the attribute MATLAB_class = "double"
  (a scalar, null-terminated ASCII string of 6 bytes).

The HDF5 C library is called through ctypes; it is searched in the loader path and in $CONDA_PREFIX/lib and
/opt/conda/lib. `available()` says whether it was found.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_hid = C.c_int64
_hsize = C.c_uint64
_herr = C.c_int

H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5T_STR_NULLTERM = 0
H5T_CSET_ASCII = 0
H5_INDEX_NAME = 0
H5_ITER_INC = 0

HEADER_TEXT = b"MATLAB 7.3 MAT-file"
USERBLOCK = 512

# MATLAB classes read as numeric arrays
_NUMERIC = {b"double", b"single", b"int8", b"uint8", b"int16", b"uint16", b"int32", b"uint32", b"int64", b"uint64",
            b"logical"}

_L = None


def _candidates():
    names = ["libhdf5.so", "libhdf5.so.103", "libhdf5.so.200", "libhdf5.so.310"]
    dirs = [None]
    if os.environ.get("CONDA_PREFIX"):
        dirs.append(Path(os.environ["CONDA_PREFIX"]) / "lib")
    dirs.append(Path("/opt/conda/lib"))
    for d in dirs:
        for n in names:
            yield n if d is None else str(d / n)


def _lib():
    global _L
    if _L is not None:
        return _L
    err = None
    for cand in _candidates():
        try:
            L = C.CDLL(cand)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"the HDF5 C library (libhdf5) was not found: {err}")
    if L.H5open() < 0:
        raise OSError("H5open failed")
    L.H5Eset_auto2(_hid(0), None, None)  # errors are reported by return codes here, not printed
    for fn, res, args in [
        ("H5Pcreate", _hid, [_hid]), ("H5Pset_userblock", _herr, [_hid, _hsize]), ("H5Pclose", _herr, [_hid]),
        ("H5Fcreate", _hid, [C.c_char_p, C.c_uint, _hid, _hid]), ("H5Fopen", _hid, [C.c_char_p, C.c_uint, _hid]),
        ("H5Fclose", _herr, [_hid]),
        ("H5Screate_simple", _hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]), ("H5Screate", _hid, [C.c_int]),
        ("H5Sclose", _herr, [_hid]), ("H5Sget_simple_extent_ndims", C.c_int, [_hid]),
        ("H5Sget_simple_extent_dims", C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
        ("H5Dcreate2", _hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        ("H5Dopen2", _hid, [_hid, C.c_char_p, _hid]), ("H5Dclose", _herr, [_hid]), ("H5Dget_space", _hid, [_hid]),
        ("H5Dwrite", _herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        ("H5Dread", _herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        ("H5Tcopy", _hid, [_hid]), ("H5Tset_size", _herr, [_hid, C.c_size_t]), ("H5Tget_size", C.c_size_t, [_hid]),
        ("H5Tset_strpad", _herr, [_hid, C.c_int]), ("H5Tset_cset", _herr, [_hid, C.c_int]), ("H5Tclose", _herr, [_hid]),
        ("H5Acreate2", _hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]), ("H5Awrite", _herr, [_hid, _hid, C.c_void_p]),
        ("H5Aopen", _hid, [_hid, C.c_char_p, _hid]), ("H5Aread", _herr, [_hid, _hid, C.c_void_p]),
        ("H5Aget_type", _hid, [_hid]), ("H5Aclose", _herr, [_hid]), ("H5Aexists", C.c_int, [_hid, C.c_char_p]),
        ("H5Lget_name_by_idx", C.c_ssize_t, [_hid, C.c_char_p, C.c_int, C.c_int, _hsize, C.c_char_p, C.c_size_t, _hid]),
    ]:
        f = getattr(L, fn)
        f.restype = res
        f.argtypes = args
    L._native_double = _hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    L._c_s1 = _hid.in_dll(L, "H5T_C_S1_g").value
    L._fcpl_class = _hid.in_dll(L, "H5P_CLS_FILE_CREATE_ID_g").value
    _L = L
    return L


def available() -> bool:
    try:
        _lib()
        return True
    except OSError:
        return False


def _check(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


def _string_type(L, n):
    t = _check(L.H5Tcopy(L._c_s1), "H5Tcopy")
    _check(L.H5Tset_size(t, n), "H5Tset_size")
    _check(L.H5Tset_strpad(t, H5T_STR_NULLTERM), "H5Tset_strpad")
    _check(L.H5Tset_cset(t, H5T_CSET_ASCII), "H5Tset_cset")
    return t


def write(filename, arrays: dict) -> None:
    """Writes each (name, array) as a float64 MATLAB variable of a MAT v7.3 file (MAT.matwrite on a
    Dict{String, Any} of Float64 arrays)."""
    L = _lib()
    filename = str(filename)
    fcpl = _check(L.H5Pcreate(L._fcpl_class), "H5Pcreate")
    try:
        _check(L.H5Pset_userblock(fcpl, USERBLOCK), "H5Pset_userblock")
        f = _check(L.H5Fcreate(filename.encode(), H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), f"creating {filename}")
    finally:
        L.H5Pclose(fcpl)
    try:
        for name, v in arrays.items():
            a = np.asfortranarray(np.asarray(v, dtype=np.float64))
            if a.size == 0:
                raise ValueError(f"{name}: empty arrays are not written")
            if a.ndim == 0:
                a = a.reshape(1, 1)
            dims = (_hsize * a.ndim)(*a.shape[::-1])
            space = _check(L.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
            try:
                d = _check(L.H5Dcreate2(f, name.encode(), L._native_double, space, H5P_DEFAULT, H5P_DEFAULT,
                                        H5P_DEFAULT), f"creating {name}")
            finally:
                L.H5Sclose(space)
            try:
                _check(L.H5Dwrite(d, L._native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), f"writing {name}")
                t = _string_type(L, 6)
                s = _check(L.H5Screate(H5S_SCALAR), "H5Screate")
                at = _check(L.H5Acreate2(d, b"MATLAB_class", t, s, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
                _check(L.H5Awrite(at, t, C.c_char_p(b"double")), "H5Awrite")
                L.H5Aclose(at)
                L.H5Sclose(s)
                L.H5Tclose(t)
            finally:
                L.H5Dclose(d)
    finally:
        _check(L.H5Fclose(f), "H5Fclose")
    header = bytearray(USERBLOCK)
    header[:len(HEADER_TEXT)] = HEADER_TEXT
    header[125], header[126], header[127] = 0x02, ord("I"), ord("M")
    with open(filename, "r+b") as fh:
        fh.write(bytes(header))


def _attr_string(L, obj, name: bytes):
    if L.H5Aexists(obj, name) <= 0:
        return None
    at = _check(L.H5Aopen(obj, name, H5P_DEFAULT), "H5Aopen")
    try:
        t = L.H5Aget_type(at)
        n = L.H5Tget_size(t)
        mt = _string_type(L, n + 1)
        buf = C.create_string_buffer(n + 1)
        _check(L.H5Aread(at, mt, buf), "H5Aread")
        L.H5Tclose(mt)
        L.H5Tclose(t)
        return buf.value
    finally:
        L.H5Aclose(at)


def read(filename) -> dict:
    """The numeric variables of a MAT v7.3 file as float64 arrays in column-major order (MAT.matread; structs, cells
    and character arrays are skipped)."""
    L = _lib()
    f = _check(L.H5Fopen(str(filename).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"opening {filename}")
    out = {}
    try:
        i = 0
        while True:
            n = L.H5Lget_name_by_idx(f, b".", H5_INDEX_NAME, H5_ITER_INC, i, None, 0, H5P_DEFAULT)
            if n < 0:
                break
            buf = C.create_string_buffer(n + 1)
            L.H5Lget_name_by_idx(f, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, n + 1, H5P_DEFAULT)
            i += 1
            name = buf.value
            d = L.H5Dopen2(f, name, H5P_DEFAULT)
            if d < 0:
                continue  # a group: struct or cell
            try:
                cls = _attr_string(L, d, b"MATLAB_class")
                if cls is not None and cls not in _NUMERIC:
                    continue
                space = L.H5Dget_space(d)
                nd = L.H5Sget_simple_extent_ndims(space)
                dims = (_hsize * max(nd, 1))()
                L.H5Sget_simple_extent_dims(space, dims, None)
                L.H5Sclose(space)
                shape = tuple(int(dims[k]) for k in range(nd))[::-1]
                if L.H5Aexists(d, b"MATLAB_empty") > 0:
                    out[name.decode()] = np.zeros((0, 0))
                    continue
                a = np.empty(shape if shape else (1,), dtype=np.float64, order="F")
                _check(L.H5Dread(d, L._native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data),
                       f"reading {name!r}")
                out[name.decode()] = a
            finally:
                L.H5Dclose(d)
    finally:
        L.H5Fclose(f)
    return out


def is_mat73(filename) -> bool:
    with open(filename, "rb") as fh:
        head = fh.read(USERBLOCK + 8)
    return head.startswith(b"MATLAB 7.3") or head[USERBLOCK:USERBLOCK + 8] == b"\x89HDF\r\n\x1a\n"
