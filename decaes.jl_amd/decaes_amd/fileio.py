"""Image and result files for the command line layer (src/main.jl:577-604 load_image, :377-411 MAT output).

* MAT: results are written as MAT v7.3 (HDF5, through the HDF5 C library; decaes_amd.mat73) with the reference's
  keys and array shapes, as MAT.jl's matwrite writes them; where libhdf5 is missing they are written as MAT v5 with
  a warning. Inputs: v7.3 files through libhdf5, v4/v5/v7 files through scipy.io.
* NIfTI-1 (.nii, .nii.gz): a reader and writer of the single-file format, with scl_slope / scl_inter applied as the
  reference does (slope 0 means unscaled).
* PAR/REC and XML/REC (.par, .xml, .rec): Philips' formats, read by decaes_amd.parrec and scaled to floating-point
  values.
"""
from __future__ import annotations

import gzip
import struct
import warnings
from pathlib import Path

import numpy as np

ALLOWED_FILE_SUFFIXES = (".mat", ".nii", ".nii.gz", ".par", ".xml", ".rec")

# NIfTI-1 datatype codes -> numpy dtypes
_NII_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8, 512: np.uint16,
               768: np.uint32, 1024: np.int64, 1280: np.uint64}
_NII_CODES = {np.dtype(v): k for k, v in _NII_DTYPES.items()}


def maybe_get_suffix(filename: str):
    """The allowed suffix `filename` ends with (case-insensitive), or None (src/main.jl maybe_get_suffix)."""
    low = str(filename).lower()
    for ext in ALLOWED_FILE_SUFFIXES:
        if low.endswith(ext):
            return ext
    return None


def chop_allowed_suffix(filename: str) -> str:
    ext = maybe_get_suffix(filename)
    return filename[: len(filename) - len(ext)] if ext else filename


def _open(path):
    return gzip.open(path, "rb") if str(path).lower().endswith(".gz") else open(path, "rb")


def read_nifti(path) -> np.ndarray:
    """NIfTI-1 single file (.nii / .nii.gz) -> array with scl_slope / scl_inter applied (Fortran order)."""
    with _open(path) as f:
        raw = f.read()
    if len(raw) < 348:
        raise ValueError(f"{path}: too short for a NIfTI-1 header")
    endian = "<" if struct.unpack("<i", raw[:4])[0] == 348 else ">"
    if struct.unpack(endian + "i", raw[:4])[0] != 348:
        raise ValueError(f"{path}: not a NIfTI-1 file (sizeof_hdr != 348)")
    dim = struct.unpack(endian + "8h", raw[40:56])
    datatype, bitpix = struct.unpack(endian + "2h", raw[70:74])
    vox_offset = struct.unpack(endian + "f", raw[108:112])[0]
    scl_slope, scl_inter = struct.unpack(endian + "2f", raw[112:120])
    if datatype not in _NII_DTYPES:
        raise ValueError(f"{path}: unsupported NIfTI datatype {datatype}")
    ndim = dim[0]
    shape = tuple(int(d) for d in dim[1:1 + ndim])
    dt = np.dtype(_NII_DTYPES[datatype]).newbyteorder(endian)
    count = int(np.prod(shape)) if shape else 1
    off = int(vox_offset) if vox_offset >= 348 else 352
    data = np.frombuffer(raw, dtype=dt, count=count, offset=off).reshape(shape, order="F")
    if scl_slope == 0 or not np.isfinite(scl_slope):
        scl_slope, scl_inter = 1.0, 0.0
    # data * scl_slope + scl_inter in FP64 into one column-major array: no full-size temporaries
    out = np.empty(shape, dtype=np.float64, order="F")
    np.multiply(data, scl_slope, out=out, dtype=np.float64)
    np.add(out, scl_inter, out=out)
    return out


def write_nifti(path, data: np.ndarray, pixdim=(1.0, 1.0, 1.0)) -> None:
    """Writes `data` as a NIfTI-1 single file (gzip when the name ends in .gz), unscaled, little endian."""
    data = np.asarray(data)
    if data.dtype not in _NII_CODES:
        data = data.astype(np.float64)
    code = _NII_CODES[data.dtype]
    dims = [data.ndim, *data.shape] + [1] * (7 - data.ndim)
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<2h", hdr, 70, code, data.dtype.itemsize * 8)
    struct.pack_into("<8f", hdr, 76, 1.0, *pixdim, 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<2f", hdr, 112, 1.0, 0.0)
    hdr[344:348] = b"n+1\x00"
    payload = bytes(hdr) + np.asfortranarray(data).astype(data.dtype.newbyteorder("<")).tobytes(order="F")
    if str(path).lower().endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(payload)
    else:
        Path(path).write_bytes(payload)


def ensure_ndims(filename, data: np.ndarray, n: int) -> np.ndarray:
    """Adds trailing singleton dimensions or selects the first volume along extra trailing ones
    (src/main.jl ensure_ndims)."""
    if data.ndim < n:
        return data.reshape(data.shape + (1,) * (n - data.ndim), order="F")
    if data.ndim > n:
        if any(s > 1 for s in data.shape[n:]):
            warnings.warn(f"Input file {filename} has {data.ndim} dimensions, expected {n}; selecting the first "
                          f"{n}-D volume along dimension(s) {', '.join(str(i + 1) for i in range(n, data.ndim))}.")
        return data[(slice(None),) * n + (0,) * (data.ndim - n)]
    return data


def load_image(filename, ndims: int = 4) -> np.ndarray:
    """load_image (src/main.jl:577-621): the first `ndims`-D array of a MAT file, a NIfTI image or a PAR/REC or
    XML/REC image; Float64, Fortran order."""
    ext = maybe_get_suffix(filename)
    if ext == ".mat":
        data = read_mat(filename)
        keys = sorted(k for k, v in data.items() if not k.startswith("__") and isinstance(v, np.ndarray)
                      and v.ndim == ndims)
        if not keys:
            raise ValueError(f"No {ndims}-D array was found in the input file: {filename}")
        if len(keys) > 1:
            warnings.warn(f"Multiple possible images found in file: {filename}\nChoosing variable {keys[0]!r} out of "
                          f"the following options: {', '.join(repr(k) for k in keys)}")
        arr = data[keys[0]]
    elif ext in (".nii", ".nii.gz"):
        arr = read_nifti(filename)
    elif ext in (".par", ".xml", ".rec"):
        from . import parrec

        arr = parrec.load(filename)
    else:
        raise ValueError(f"Currently, only {', '.join(ALLOWED_FILE_SUFFIXES)} files are supported")
    arr = ensure_ndims(filename, np.asarray(arr), ndims)
    return np.asfortranarray(arr, dtype=np.float64)


def read_mat(filename) -> dict:
    """The variables of a MAT file (v7.3 through libhdf5, older versions through scipy.io)."""
    from . import mat73

    if mat73.is_mat73(filename):
        return mat73.read(filename)
    import scipy.io

    return scipy.io.loadmat(filename)


def save_mat(filename, arrays: dict) -> None:
    """MAT file with one float64 variable per key, arrays keeping their shapes: v7.3 as MAT.matwrite writes it
    (src/main.jl:382,391,411), or v5 when the HDF5 library is missing."""
    from . import mat73

    if mat73.available():
        mat73.write(filename, arrays)
        return
    import scipy.io

    warnings.warn("libhdf5 not found: writing MAT v5 instead of v7.3")
    scipy.io.savemat(filename, {k: (np.asfortranarray(v) if isinstance(v, np.ndarray) else v)
                                for k, v in arrays.items()}, oned_as="column")
