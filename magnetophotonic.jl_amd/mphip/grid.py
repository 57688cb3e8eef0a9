"""Grid axes, physical constants and the CFL step.

Host-side preprocessing restated from src/grid/Grid.jl and src/core/Constants.jl of the
reference.  Everything here is FP64 numpy on the host; its outputs (the 1-D inv_d_cell /
inv_d_dual arrays and dt) are inputs to the device kernels.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

# src/core/Constants.jl:13-24 (FDTDParams defaults).  4pi*1e-7 and 299792458.0^2 are
# evaluated exactly as Julia does (4*pi in Float64, x^2 as x*x).
C0 = 299792458.0
MU0 = 4 * math.pi * 1e-7
EPS0 = 1.0 / ((4 * math.pi * 1e-7) * (299792458.0 * 299792458.0))


@dataclass(frozen=True)
class FDTDParams:
    """FDTDParams{Float64} (src/core/Constants.jl:3-25)."""
    c0: float = C0
    mu0: float = MU0
    eps0: float = EPS0
    n_si3n4: float = 2.0
    epsr_si3n4: float = 4.0
    n_sio2: float = 1.444
    epsr_sio2: float = 1.444 * 1.444
    epsr_vac: float = 1.0


def get_n_sio2(lambda_m: float) -> float:
    """Sellmeier SiO2 (src/core/Constants.jl:27-35)."""
    l_um = float(lambda_m) * 1e6
    l2 = l_um * l_um
    n2 = (1.0 + (0.6961663 * l2) / (l2 - 0.00467914826) + (0.4079426 * l2) / (l2 - 0.0135120631)
          + (0.8974794 * l2) / (l2 - 97.9340025))
    return math.sqrt(n2)


def get_n_si3n4(lambda_m: float) -> float:
    """Sellmeier Si3N4 (src/core/Constants.jl:37-44)."""
    l_um = float(lambda_m) * 1e6
    l2 = l_um * l_um
    n2 = 1.0 + (3.0249 * l2) / (l2 - 0.13534 * 0.13534) + (40314.0 * l2) / (l2 - 1239.842 * 1239.842)
    return math.sqrt(n2)


def cfl_dt_spacing(dx: float, dy: float, dz: float, p: FDTDParams = FDTDParams(),
                   courant: float = 0.99) -> float:
    """cfl_dt(dx, dy, dz; courant) (src/core/Constants.jl:61-64)."""
    ix, iy, iz = 1.0 / dx, 1.0 / dy, 1.0 / dz
    denom = p.c0 * math.sqrt(ix * ix + iy * iy + iz * iz)
    return float(courant) / denom


@dataclass
class Axis1D:
    """Axis1D (src/grid/Grid.jl:1-7)."""
    edges: np.ndarray
    centers: np.ndarray
    inv_d_cell: np.ndarray
    inv_d_dual: np.ndarray
    d_min: float

    @property
    def n(self) -> int:
        return len(self.centers)

    def widths(self) -> np.ndarray:
        return np.diff(self.edges)


@dataclass
class Grid3D:
    x: Axis1D
    y: Axis1D
    z: Axis1D

    @property
    def shape(self):
        return (self.x.n, self.y.n, self.z.n)

    @property
    def ncells(self) -> int:
        return self.x.n * self.y.n * self.z.n


def _axis_from_edges(edges, dual_in_dmin: bool = True) -> Axis1D:
    """_axis_from_edges / _axis_from_ref_edges (src/grid/Grid.jl:24-38, 224-236)."""
    e = np.asarray(edges, dtype=np.float64)
    if len(e) < 2:
        raise ValueError("axis requires at least two edges")
    widths = np.diff(e)
    if not np.all(widths > 0.0):
        raise ValueError("axis edges must be strictly increasing")
    n = len(e) - 1
    centers = (e[:-1] + e[1:]) / 2.0
    inv_d_cell = 1.0 / widths
    inv_d_dual = np.zeros(n)
    inv_d_dual[0] = 1.0 / (centers[0] - e[0])
    inv_d_dual[1:] = 1.0 / (centers[1:] - centers[:-1])
    dmin = min(float(widths.min()), 1.0 / float(inv_d_dual.max())) if dual_in_dmin else float(widths.min())
    return Axis1D(e, centers, inv_d_cell, inv_d_dual, dmin)


def uniform_axis(xmin: float, xmax: float, dx: float) -> Axis1D:
    """uniform_axis (src/grid/Grid.jl:54-60).  Edges are the correctly rounded values of
    xmin + i*(xmax-xmin)/n, which is what Julia's Float64 range(...; length) returns."""
    if not xmax > xmin:
        raise ValueError("xmax must exceed xmin")
    if not dx > 0:
        raise ValueError("dx must be positive")
    n = max(1, int(round((float(xmax) - float(xmin)) / float(dx))))
    a, b = Fraction(float(xmin)), Fraction(float(xmax))
    edges = [float(a + (b - a) * i / n) for i in range(n + 1)]
    return _axis_from_edges(edges)


def uniform_axis_n(n: int, dx: float, xmin: float = 0.0) -> Axis1D:
    """An n-cell uniform axis of spacing dx starting at xmin."""
    return uniform_axis(xmin, xmin + n * dx, dx)


def _fill_linear_region(edges: list, start_val: float, end_val: float, target_d: float) -> float:
    """_fill_linear_region! (src/grid/Grid.jl:62-75)."""
    if end_val <= start_val + 1e-12:
        return start_val
    dist = end_val - start_val
    n_cells = max(1, int(round(dist / target_d)))
    actual_d = dist / n_cells
    curr = start_val
    for _ in range(n_cells):
        curr += actual_d
        edges.append(curr)
    return curr


def _dedup_sorted_edges(edges: list, tol: float) -> list:
    """_dedup_sorted_edges (src/grid/Grid.jl:44-52)."""
    edges = sorted(edges)
    out = [edges[0]]
    for v in edges[1:]:
        if v - out[-1] > tol:
            out.append(v)
    out[-1] = edges[-1]
    return out


def graded_axis(L_phys: float, d_fine: float, fine_half_width: float, center: float,
                d_max: float, stretch_ratio: float) -> Axis1D:
    """graded_axis (src/grid/Grid.jl:77-95): fine cells of d_fine within
    +-fine_half_width of center, geometric growth up to d_max outside."""
    L, c = float(L_phys), float(center)
    edges_right = [c]
    curr = _fill_linear_region(edges_right, c, c + float(fine_half_width), float(d_fine))
    curr_d = float(d_fine)
    while curr < c + L / 2.0:
        curr_d = min(curr_d * float(stretch_ratio), float(d_max))
        curr += curr_d
        edges_right.append(curr)
    edges_right[-1] = c + L / 2.0
    edges_left = [c - (edges_right[i] - c) for i in range(len(edges_right) - 1, 0, -1)]
    return _axis_from_edges(_dedup_sorted_edges(edges_left + edges_right, 1e-3 * float(d_fine)))


def shift_axis(axis: Axis1D, dx: float) -> Axis1D:
    """_shift_axis (src/grid/Grid.jl:158-160)."""
    return _axis_from_edges(axis.edges + float(dx))


def ref_graded_axis(L_phys, d_fine, fine_half_width, center, d_max, stretch_ratio) -> Axis1D:
    """ref_graded_axis (src/grid/Grid.jl:238-254)."""
    L, c, df, dm, sr = float(L_phys), float(center), float(d_fine), float(d_max), float(stretch_ratio)
    edges_right = [c]
    curr = _fill_linear_region(edges_right, c, c + float(fine_half_width), df)
    curr_d = df
    while curr < c + L / 2.0:
        curr_d = min(curr_d * sr, dm)
        curr += curr_d
        edges_right.append(curr)
    edges_right[-1] = c + L / 2.0
    edges_left = [c - (edges_right[i] - c) for i in range(len(edges_right) - 1, 0, -1)]
    return _axis_from_edges(edges_left + edges_right, dual_in_dmin=True)


def ref_propagation_axis(L_phys, d_coarse, d_fine, x_film_start, x_film_end, buffer,
                         stretch_ratio) -> Axis1D:
    """ref_propagation_axis (src/grid/Grid.jl:256-288)."""
    L, dc, df, sr = float(L_phys), float(d_coarse), float(d_fine), float(stretch_ratio)
    x_fine_start = float(x_film_start) - float(buffer)
    x_fine_end = float(x_film_end) + float(buffer)
    edges = [0.0]
    grade_cells = []
    d_curr = dc
    while d_curr > df * 1.001:
        d_curr = max(d_curr / sr, df)
        grade_cells.append(d_curr)
    x_grade_start = x_fine_start - _julia_sum(grade_cells)
    curr = _fill_linear_region(edges, 0.0, x_grade_start, dc)
    for gcell in grade_cells:
        curr += gcell
        edges.append(curr)
    edges[-1] = x_fine_start
    curr = x_fine_start
    curr = _fill_linear_region(edges, curr, float(x_film_start), df)
    curr = _fill_linear_region(edges, curr, float(x_film_end), df)
    curr = _fill_linear_region(edges, curr, x_fine_end, df)
    d_curr = df
    while d_curr < dc * 0.999:
        d_curr = min(d_curr * sr, dc)
        curr += d_curr
        edges.append(curr)
    curr = _fill_linear_region(edges, curr, L, dc)
    edges[-1] = L
    edges = sorted(set(edges))
    return _axis_from_edges(edges, dual_in_dmin=False)


def _julia_sum(v) -> float:
    """Base.sum of a Vector{Float64}: mapreduce with pairwise blocking (block size 1024);
    below the block size it is the 4-way-unrolled SIMD reduction of mapreduce_impl."""
    v = [float(x) for x in v]
    n = len(v)
    if n == 0:
        return 0.0
    if n < 16:
        s = v[0]
        for x in v[1:]:
            s += x
        return s
    # reduce_impl for n >= 16 and n <= 1024: four interleaved accumulators
    a1, a2, a3, a4 = v[0], v[1], v[2], v[3]
    i = 4
    while i + 3 < n:
        a1 += v[i]
        a2 += v[i + 1]
        a3 += v[i + 2]
        a4 += v[i + 3]
        i += 4
    s = (a1 + a2) + (a3 + a4)
    while i < n:
        s += v[i]
        i += 1
    return s


def not_gate_reference_grid() -> Grid3D:
    """The validated 4099x190x100 NOT-gate mesh (src/grid/Grid.jl:295-300)."""
    x = ref_propagation_axis(60.0e-6, 15e-9, 4e-9, 40.0e-6, 40.008e-6, 200e-9, 1.05)
    y = ref_graded_axis(6.0e-6, 20e-9, 1.2e-6, 0.0, 150e-9, 1.05)
    z = ref_graded_axis(3.0e-6, 20e-9, 0.5e-6, 0.2e-6, 120e-9, 1.05)
    return Grid3D(x, y, z)


def cfl_dt(grid: Grid3D, p: FDTDParams = FDTDParams(), courant: float = 0.99) -> float:
    """cfl_dt(grid::Grid3D; courant) (src/grid/Grid.jl:350-352): uses each axis' d_min."""
    return cfl_dt_spacing(grid.x.d_min, grid.y.d_min, grid.z.d_min, p, courant)


def cfl_dt_cells(grid: Grid3D, p: FDTDParams = FDTDParams(), courant: float = 0.99) -> float:
    """cfl_dt(dx, dy, dz) on the smallest primary cell width of each axis (dual cells
    excluded): the step used for the synthetic workloads."""
    return cfl_dt_spacing(float(grid.x.widths().min()), float(grid.y.widths().min()),
                          float(grid.z.widths().min()), p, courant)
