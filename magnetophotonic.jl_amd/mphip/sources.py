"""Source waveforms and placement (host side).

Restates src/fdtd/Source.jl: the pulse values evaluated on the host in FP64 once per
step (:1-33), the nearest-centre index (:63-72), the two-cell interpolation weights of a
plane source (:74-86) and the source kinds (:35-61).  Injection itself runs on the device
(mp_inject_soft_3d / mp_inject_mode_x_3d / mp_inject_plane_3d).
"""
from __future__ import annotations

import bisect
import math
from dataclasses import dataclass
from typing import Any

import numpy as np

from .grid import Axis1D


@dataclass
class GaussianPulse:
    """GaussianPulse (Source.jl:1-10): A exp(-((t-t0)/tau)^2) sin(w(t-t0)+phase)."""
    amplitude: float = 1.0
    t0: float = 160e-15
    tau: float = 40e-15
    omega: float = 2 * math.pi * 299792458.0 / 800e-9
    phase: float = 0.0

    def value(self, t: float) -> float:
        u = (float(t) - self.t0) / self.tau
        return self.amplitude * math.exp(-(u * u)) * math.sin(self.omega * (float(t) - self.t0) + self.phase)


@dataclass
class CarrierGaussianPulse:
    """CarrierGaussianPulse (Source.jl:22-30): A sin(w t) exp(-1/2 ((t-t0)/sigma)^2)."""
    amplitude: float = 1.0
    t0: float = 160e-15
    sigma: float = 40e-15
    omega: float = 2 * math.pi * 299792458.0 / 800e-9

    def value(self, t: float) -> float:
        u = (float(t) - self.t0) / self.sigma
        return self.amplitude * math.sin(self.omega * float(t)) * math.exp(-0.5 * (u * u))


@dataclass
class ContinuousSource:
    """ContinuousSource (Source.jl:12-16, 33)."""
    amplitude: float = 1.0
    omega: float = 2 * math.pi * 299792458.0 / 800e-9
    phase: float = 0.0

    def value(self, t: float) -> float:
        return self.amplitude * math.sin(self.omega * float(t) + self.phase)


def source_value(pulse, t: float) -> float:
    """source_value (Source.jl:29-33)."""
    return pulse.value(t)


COMPONENT_CODE = {"Ex": 1, "Ey": 2, "Ez": 3, "Hx": 4, "Hy": 5, "Hz": 6}
AXIS_CODE = {"x": 0, "y": 1, "z": 2}


@dataclass
class SoftSource:
    """The (pulse, component, index) tuple source of FDTDState (Solver.jl:296-302);
    index is 0-based (i, j, k)."""
    pulse: Any
    component: str
    index: tuple


@dataclass
class PlaneSource:
    """PlaneSource (Source.jl:43-48, 59)."""
    pulse: Any
    component: str
    axis: str = "x"
    position: float = 0.0


@dataclass
class ModeSource:
    """ModeSource (Source.jl:50-61); profile is (ny, nz) for axis x."""
    profile: np.ndarray
    pulse: Any
    neff: float = 1.0
    axis: str = "x"
    position: Any = 0.0
    component: str = "Ez"


def nearest_index(axis: Axis1D, x: float) -> int:
    """_nearest_index (Source.jl:65-70), returned 0-based; ties pick the lower index."""
    c = axis.centers
    n = len(c)
    hi = min(max(bisect.bisect_left(list(c), float(x)) + 1, 1), n)  # searchsortedfirst, 1-based
    lo = max(hi - 1, 1)
    pick = lo if abs(c[lo - 1] - float(x)) <= abs(c[hi - 1] - float(x)) else hi
    return pick - 1


def source_index(axis: Axis1D, pos) -> int:
    """_source_index (Source.jl:71-72): an integer is a (0-based) index, a float a position."""
    if isinstance(pos, (int, np.integer)):
        return min(max(int(pos), 0), axis.n - 1)
    return nearest_index(axis, float(pos))


def axis_weights(axis: Axis1D, x: float):
    """_axis_weights (Source.jl:74-86): ((lo, w_lo), (hi, w_hi)), 0-based."""
    xf = float(x)
    c = axis.centers
    n = len(c)
    if xf <= c[0]:
        return ((0, 1.0),)
    if xf >= c[-1]:
        return ((n - 1, 1.0),)
    hi = bisect.bisect_left(list(c), xf)  # 0-based searchsortedfirst
    lo = hi - 1
    w_hi = (xf - c[lo]) / (c[hi] - c[lo])
    return ((lo, 1.0 - w_hi), (hi, w_hi))


def gaussian_mode_profile(ny: int, nz: int, dy: float, dz: float, waist_y=None,
                          waist_z=None) -> np.ndarray:
    """gaussian_mode_profile + _normalize_mode (src/fdtd/ModeSolver.jl:1-19): the Gaussian
    the mode source uses without an eigenmode solve, normalised to sum(p^2) dy dz = 1."""
    wy = max(2 * dy, 0.25 * ny * dy) if waist_y is None else float(waist_y)
    wz = max(2 * dz, 0.25 * nz * dz) if waist_z is None else float(waist_z)
    y = (np.arange(1, ny + 1) - 0.5 * (ny + 1)) * dy
    z = (np.arange(1, nz + 1) - 0.5 * (nz + 1)) * dz
    prof = np.exp(-(y[:, None] / wy) ** 2 - (z[None, :] / wz) ** 2)
    norm = math.sqrt(float(np.sum(prof * prof)) * dy * dz)
    return prof / norm if norm > 0 else prof
