"""decaes_amd: MI355X (gfx950) backend of DECAES.jl's T2mapSEcorr / T2partSEcorr voxelwise path.

Public API mirrors the reference (src/DECAES.jl:53): T2mapOptions, T2mapSEcorr, T2partOptions, T2partSEcorr.
The numerical work runs in libdecaes_amd.so (HIP kernels behind a C ABI, include/decaes_amd.h).
"""
from .options import T2mapOptions, T2partOptions
from .api import T2mapSEcorr, T2partSEcorr, decaybasisset, synth_mse
from ._lib import Context, DmiError, context, load as load_library, LIB_PATH, EXPORTS

__all__ = ["T2mapOptions", "T2partOptions", "T2mapSEcorr", "T2partSEcorr", "decaybasisset", "synth_mse",
           "Context", "DmiError", "context", "load_library", "LIB_PATH", "EXPORTS"]
__version__ = "0.1.0"
