"""T2mapOptions / T2partOptions: the reference's option structs with the same fields, defaults and checks.

Mirrors src/types.jl:19-100 (T2mapOptions) and :139-190 (T2partOptions). Failed checks raise
AssertionError with the reference's @assert messages. `precision` is this backend's one extra field:
"fp64" (default, the plain FP64 path) or "mixed" (FP32 EPG bases, FP64 NNLS and regularization).
"""
from __future__ import annotations

import math
import os
from fractions import Fraction
from dataclasses import dataclass, field, fields
from typing import Optional, Tuple

import numpy as np


def _linrange(a: float, b: float, n: int) -> np.ndarray:
    """range(a, b; length = n) with Julia's correctly rounded elements and exact endpoints."""
    if n == 1:
        return np.array([float(a)])
    fa, fb = Fraction(a), Fraction(b)
    r = np.array([float(((n - 1 - j) * fa + j * fb) / (n - 1)) for j in range(n)])
    r[0], r[-1] = a, b
    return r


def _logrange(a: float, b: float, n: int) -> np.ndarray:
    """logrange (src/utils.jl:7): exp of a linear range in log space, exact endpoints."""
    r = np.array([math.exp(v) for v in _linrange(math.log(a), math.log(b), n)])  # libm exp, as the library
    r[0], r[-1] = a, b
    return r


@dataclass
class T2mapOptions:
    MatrixSize: Tuple[int, int, int]
    nTE: int
    TE: float
    nT2: int
    T2Range: Tuple[float, float]
    Reg: str
    legacy: bool = False
    Threaded: bool = field(default_factory=lambda: (os.cpu_count() or 1) > 1)
    T1: float = 1.0
    Threshold: float = 0.0
    MinRefAngle: float = 50.0
    nRefAngles: Optional[int] = None
    nRefAnglesMin: Optional[int] = None
    Chi2Factor: Optional[float] = None
    NoiseLevel: Optional[float] = None
    RefConAngle: float = 180.0
    SetFlipAngle: Optional[float] = None
    SaveResidualNorm: bool = False
    SaveDecayCurve: bool = False
    SaveRegParam: bool = False
    SaveNNLSBasis: bool = False
    Silent: bool = False
    precision: str = "fp64"
    schedule: str = "auto"   # voxel-kernel launch schedule: "auto", "lds" or "split" (results are identical)

    def __post_init__(self):
        self.MatrixSize = tuple(int(v) for v in self.MatrixSize)
        self.T2Range = tuple(float(v) for v in self.T2Range)
        if self.nRefAngles is None:
            self.nRefAngles = 64 if not self.legacy else 8
        if self.nRefAnglesMin is None:
            self.nRefAnglesMin = min(5, self.nRefAngles) if not self.legacy else self.nRefAngles
        assert all(v >= 1 for v in self.MatrixSize), \
            f"MatrixSize must be a tuple of 3 positive integers, but MatrixSize = {self.MatrixSize}."
        assert self.nTE >= 4, f"At least four echoes are required for T2 mapping, but nTE = {self.nTE}."
        assert self.TE > 0.0, f"Echo spacing must be positive, but TE = {self.TE}."
        assert self.nT2 >= 2, f"At least two T2 components are required for T2 mapping, but nT2 = {self.nT2}."
        assert 0.0 < self.T2Range[0] < self.T2Range[1], \
            f"T2Range must a sorted 2-tuple of positive values, but T2Range = {self.T2Range}."
        assert self.T1 > 0.0, f"T1 must be positive, but T1 = {self.T1}."
        assert self.Threshold >= 0.0 or self.Threshold == -math.inf, (
            "First echo signal threshold must be non-negative or -Inf to force processing of every voxel, "
            f"but Threshold = {self.Threshold}.")
        assert 0.0 <= self.MinRefAngle <= 180.0, \
            f"Minimum refocusing angle must be in the range [0, 180], but MinRefAngle = {self.MinRefAngle}."
        assert self.nRefAngles >= 2, (
            "Maximum number of angles to check during flip angle optimization must be at least 2, "
            f"but nRefAngles = {self.nRefAngles}.")
        assert 2 <= self.nRefAnglesMin <= self.nRefAngles, (
            "Minimum number of angles to check during flip angle optimization must be in the range [2, nRefAngles], "
            f"but nRefAngles = {self.nRefAngles} and nRefAnglesMin = {self.nRefAnglesMin}.")
        assert self.Reg in ("none", "lcurve", "gcv", "chi2", "mdp")
        assert self.Reg != "chi2" or (self.Chi2Factor is not None and self.Chi2Factor > 1.0), \
            f"Chi2Factor must be greater than 1.0, but Chi2Factor = {self.Chi2Factor}."
        assert self.Reg != "mdp" or (self.NoiseLevel is not None and self.NoiseLevel > 0.0), \
            f"Noise level must be positive, but NoiseLevel = {self.NoiseLevel}."
        assert 0.0 <= self.RefConAngle <= 180.0, \
            f"Refocusing control angle must be in the range [0, 180], but RefConAngle = {self.RefConAngle}."
        assert self.SetFlipAngle is None or 0.0 <= self.SetFlipAngle <= 180.0, \
            f"Fixed flip angle must be in the range [0, 180], but SetFlipAngle = {self.SetFlipAngle}."
        assert self.precision in ("fp64", "mixed"), f"precision must be 'fp64' or 'mixed', got {self.precision!r}"
        assert self.schedule in ("auto", "lds", "split"), f"schedule must be 'auto', 'lds' or 'split', got {self.schedule!r}"

    @classmethod
    def from_image(cls, image, **kwargs) -> "T2mapOptions":
        """T2mapOptions(image; kwargs...) (src/types.jl:102): MatrixSize and nTE from the image shape."""
        shape = tuple(image.shape)
        assert len(shape) == 4, f"image must be 4D (row, column, slice, echo), got shape {shape}"
        return cls(**{**kwargs, "MatrixSize": shape[:3], "nTE": shape[3]})

    def T2_component_times(self) -> np.ndarray:  # src/types.jl:108
        return _logrange(self.T2Range[0], self.T2Range[1], self.nT2)

    def flip_angles(self) -> np.ndarray:  # src/types.jl:109
        if self.SetFlipAngle is None:
            return _linrange(self.MinRefAngle, 180.0, self.nRefAngles)
        return np.array([float(self.SetFlipAngle)])

    def to_dict(self) -> dict:
        return {f.name: getattr(self, f.name) for f in fields(self)}


@dataclass
class T2partOptions:
    MatrixSize: Tuple[int, int, int]
    nT2: int
    T2Range: Tuple[float, float]
    SPWin: Tuple[float, float]
    MPWin: Tuple[float, float]
    legacy: bool = False
    Threaded: bool = field(default_factory=lambda: (os.cpu_count() or 1) > 1)
    Sigmoid: Optional[float] = None
    Silent: bool = False

    def __post_init__(self):
        self.MatrixSize = tuple(int(v) for v in self.MatrixSize)
        self.T2Range = tuple(float(v) for v in self.T2Range)
        self.SPWin = tuple(float(v) for v in self.SPWin)
        self.MPWin = tuple(float(v) for v in self.MPWin)
        assert all(v >= 1 for v in self.MatrixSize)
        assert self.nT2 >= 2
        assert 0.0 < self.T2Range[0] < self.T2Range[1]
        assert self.SPWin[0] < self.SPWin[1]
        assert self.MPWin[0] < self.MPWin[1]
        assert self.Sigmoid is None or self.Sigmoid > 0

    @classmethod
    def from_dist(cls, dist, **kwargs) -> "T2partOptions":
        """T2partOptions(t2dist; kwargs...) (src/types.jl:174)."""
        shape = tuple(dist.shape)
        assert len(shape) == 4, f"dist must be 4D (row, column, slice, T2), got shape {shape}"
        return cls(**{**kwargs, "MatrixSize": shape[:3], "nT2": shape[3]})

    @classmethod
    def from_t2map(cls, o: T2mapOptions, **kwargs) -> "T2partOptions":
        """T2partOptions(o::T2mapOptions; kwargs...) (src/types.jl:180-190)."""
        base = dict(legacy=o.legacy, Threaded=o.Threaded, MatrixSize=o.MatrixSize, nT2=o.nT2, T2Range=o.T2Range,
                    Silent=o.Silent)
        return cls(**{**base, **kwargs})
