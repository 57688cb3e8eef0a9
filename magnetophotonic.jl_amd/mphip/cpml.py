"""CFS-CPML coefficients and slab geometry (host side).

Restates src/fdtd/CPML.jl: the per-axis (inv_kappa, a, b) vectors on non-uniform edges
(`_cpml_axis`, :87-124), the slab widths (`_cpml_width`, :151) and the 24 psi slab shapes
(`build_cpml`, :172-189).  The device holds the slabs; this module only builds the
coefficients (FP64 math, rounded to the compute type exactly where the reference rounds).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import Axis1D, FDTDParams, Grid3D

# psi slab order used by the C-ABI (mp_cpml.psi_h / psi_d): psi_<component><deriv axis>_<side>
PSI_NAMES = ("xy_lo", "xy_hi", "xz_lo", "xz_hi", "yz_lo", "yz_hi",
             "yx_lo", "yx_hi", "zx_lo", "zx_hi", "zy_lo", "zy_hi")
# derivative axis of each psi slab: y,y,z,z,z,z,x,x,x,x,y,y
PSI_AXIS = (1, 1, 2, 2, 2, 2, 0, 0, 0, 0, 1, 1)


@dataclass
class CPMLAxis:
    inv_kappa: np.ndarray
    a: np.ndarray
    b: np.ndarray


def cpml_width(n: int, npml: int) -> int:
    """_cpml_width (src/fdtd/CPML.jl:151)."""
    return int(npml) if (npml > 0 and n >= 2 * npml + 1) else 0


def pml_slab(idx: int, n: int, npml: int):
    """_pml_slab (src/fdtd/Kernels.jl:9-14), 1-based idx -> (slab index, region)."""
    if npml <= 0:
        return 0, 0
    if idx <= npml:
        return idx, 1
    if idx > n - npml:
        return idx - (n - npml), 2
    return 0, 0


def cpml_axis(edges: np.ndarray, n_pml: int, dt: float, p: FDTDParams = FDTDParams(),
              dtype=np.float64, order: float = 3.0, reflection: float = 1e-8,
              kappa_max: float = 5.0, alpha_max: float = 0.05) -> CPMLAxis:
    """_cpml_axis (src/fdtd/CPML.jl:87-124).  `a` carries no 1/dx: it multiplies the
    already spacing-scaled derivative.  `a` is formed from the dtype-rounded `b`, as the
    reference does (Float64(b[i]) after b[i] = T(...))."""
    edges = np.asarray(edges, dtype=np.float64)
    n = len(edges) - 1
    inv_kappa = np.ones(n, dtype=dtype)
    a = np.zeros(n, dtype=dtype)
    b = np.zeros(n, dtype=dtype)
    if n_pml <= 0 or n < 2 * n_pml + 1:
        return CPMLAxis(inv_kappa, a, b)
    eta = math.sqrt(p.mu0 / p.eps0)
    th_lo = edges[n_pml] - edges[0]
    th_hi = edges[n] - edges[n - n_pml]
    smax_lo = -(order + 1.0) * math.log(reflection) / (2.0 * eta * th_lo)
    smax_hi = -(order + 1.0) * math.log(reflection) / (2.0 * eta * th_hi)
    for i in range(1, n + 1):
        x_c = (edges[i - 1] + edges[i]) / 2.0
        dist = 0.0
        smax = 0.0
        if i <= n_pml:
            dist = (edges[n_pml] - x_c) / th_lo
            smax = smax_lo
        elif i > n - n_pml:
            dist = (x_c - edges[n - n_pml]) / th_hi
            smax = smax_hi
        if dist > 0.0:
            sigma = smax * dist ** order
            kappa = 1.0 + (kappa_max - 1.0) * dist ** order
            alpha = alpha_max * (1.0 - dist) ** order
            inv_kappa[i - 1] = 1.0 / kappa
            b[i - 1] = math.exp(-(sigma / (kappa * p.eps0) + alpha / p.eps0) * dt)
            denom = sigma * kappa + kappa * kappa * alpha
            a[i - 1] = 0.0 if denom == 0.0 else (sigma / denom) * (float(b[i - 1]) - 1.0)
    return CPMLAxis(inv_kappa, a, b)


def build_cpml_profiles(n: int, n_pml: int, dx: float, dt: float, p: FDTDParams = FDTDParams(),
                        order=3.0, reflection=1e-8, kappa_max=1.0, alpha_max=0.05):
    """Legacy uniform profiles (src/fdtd/CPML.jl:13-31); kept for its known-answer test."""
    eta = math.sqrt(p.mu0 / p.eps0)
    smax = -(order + 1.0) * math.log(reflection) / (2.0 * eta * (n_pml * float(dx)))
    kappa = np.ones(n)
    a = np.zeros(n)
    b = np.zeros(n)
    for i in range(1, n + 1):
        if i <= n_pml:
            dist = (n_pml - i + 1.0) / n_pml
        elif i > n - n_pml:
            dist = (i - (n - n_pml)) / n_pml
        else:
            dist = 0.0
        if dist > 0.0:
            sigma = smax * dist ** order
            kappa[i - 1] = 1.0 + (kappa_max - 1.0) * dist ** order
            alpha = alpha_max * (1.0 - dist) ** order
            b[i - 1] = math.exp(-(sigma / (kappa[i - 1] * p.eps0) + alpha / p.eps0) * float(dt))
            denom = sigma * kappa[i - 1] + kappa[i - 1] ** 2 * alpha
            a[i - 1] = 0.0 if denom == 0.0 else (sigma / denom) * (b[i - 1] - 1.0) / float(dx)
    return kappa, a, b


@dataclass
class CPMLSpec:
    """Host description of a CPMLState: coefficients per axis and slab shapes."""
    axes: tuple  # (CPMLAxis x, y, z)
    widths: tuple  # (wx, wy, wz)
    shape: tuple   # (nx, ny, nz)

    def slab_shape(self, deriv_axis: int):
        nx, ny, nz = self.shape
        wx, wy, wz = self.widths
        return ((wx, ny, nz), (nx, wy, nz), (nx, ny, wz))[deriv_axis]

    def slab_cells(self, deriv_axis: int) -> int:
        s = self.slab_shape(deriv_axis)
        return int(s[0] * s[1] * s[2])


def build_cpml(grid: Grid3D, n_pml, dt: float, p: FDTDParams = FDTDParams(), dtype=np.float64,
               order=3.0, reflection=1e-8, kappa_max=5.0, alpha_max=0.05) -> CPMLSpec:
    """build_cpml (src/fdtd/CPML.jl:172-189) without the slab allocation."""
    if isinstance(n_pml, (tuple, list)):
        npx, npy, npz = (int(v) for v in n_pml)
    else:
        npx = npy = npz = int(n_pml)
    kw = dict(dtype=dtype, order=order, reflection=reflection, kappa_max=kappa_max,
              alpha_max=alpha_max)
    ax = cpml_axis(grid.x.edges, npx, dt, p, **kw)
    ay = cpml_axis(grid.y.edges, npy, dt, p, **kw)
    az = cpml_axis(grid.z.edges, npz, dt, p, **kw)
    nx, ny, nz = grid.shape
    widths = (cpml_width(nx, npx), cpml_width(ny, npy), cpml_width(nz, npz))
    return CPMLSpec((ax, ay, az), widths, (nx, ny, nz))
