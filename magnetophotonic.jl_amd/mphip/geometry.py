"""Scene rasterisation subset (host side, one-time setup).

Restates the part of src/geometry/Scene.jl:135-258 (`rasterize`) the synthetic workloads
need: axis-aligned boxes (centre-sampled in x/y at subpixel = 1, z overlap fraction,
Scene.jl:33-47, 58-62) painted in order with eps <- (1-f) eps + f eps_m, and the
magneto-optic "model" cell list keyed by linear index with the largest fill kept
(Scene.jl:182-198, 207-222).  A sphere primitive (cell-centre membership) is added for
the dispersive-sphere workload; the reference has none.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .grid import Grid3D


@dataclass
class Material:
    name: str
    epsr: float = 1.0
    dispersive: bool = False   # carries Drude-Lorentz poles (reference Material.poles non-empty)
    model: bool = False        # carries the magneto-optic film model (Material.model)


@dataclass
class Box:
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    zmin: float
    zmax: float


@dataclass
class Sphere:
    cx: float
    cy: float
    cz: float
    radius: float


@dataclass
class Scene:
    entries: list = field(default_factory=list)

    def add(self, shape, material: Material) -> "Scene":
        self.entries.append((shape, material))
        return self


@dataclass
class Geometry:
    """The NamedTuple `rasterize` returns, restricted to what FDTDState reads.
    Volumes are (nx, ny, nz) Fortran-ordered; cell lists are 0-based linear indices."""
    epsr: np.ndarray
    inv_eps: np.ndarray            # relative 1/eps_r (Scene.jl:205)
    material_cells: np.ndarray
    material_fill: np.ndarray
    material_inv_eps: np.ndarray
    pole_cells: np.ndarray
    pole_fill: np.ndarray
    pole_inv_eps: np.ndarray

    @property
    def n_material(self) -> int:
        return int(len(self.material_cells))


def _box_fraction(shape: Box, grid: Grid3D):
    """(xy membership, z fraction) of a Box at subpixel = 1 (Scene.jl:33-36, 58-62)."""
    xc, yc = grid.x.centers, grid.y.centers
    inx = (shape.xmin <= xc) & (xc <= shape.xmax)
    iny = (shape.ymin <= yc) & (yc <= shape.ymax)
    zlo, zhi = grid.z.edges[:-1], grid.z.edges[1:]
    overlap = np.minimum(zhi, shape.zmax) - np.maximum(zlo, shape.zmin)
    zf = np.clip(overlap / (zhi - zlo), 0.0, 1.0)
    return inx, iny, zf


def rasterize(scene: Scene, grid: Grid3D, background: float = 1.0) -> Geometry:
    nx, ny, nz = grid.shape
    epsr = np.full((nx, ny, nz), float(background), order="F")
    model_parts, pole_parts = [], []
    for shape, mat in scene.entries:
        if isinstance(shape, Box):
            inx, iny, zf = _box_fraction(shape, grid)
            ii = np.nonzero(inx)[0]
            jj = np.nonzero(iny)[0]
            kk = np.nonzero(zf > 0.0)[0]
            if len(ii) == 0 or len(jj) == 0 or len(kk) == 0:
                continue
            i0, i1, j0, j1, k0, k1 = ii[0], ii[-1] + 1, jj[0], jj[-1] + 1, kk[0], kk[-1] + 1
            f = (inx[i0:i1, None, None] & iny[None, j0:j1, None]).astype(np.float64) * zf[None, None, k0:k1]
        elif isinstance(shape, Sphere):
            xc, yc, zc = grid.x.centers, grid.y.centers, grid.z.centers
            r = shape.radius
            ii = np.nonzero(np.abs(xc - shape.cx) <= r)[0]
            jj = np.nonzero(np.abs(yc - shape.cy) <= r)[0]
            kk = np.nonzero(np.abs(zc - shape.cz) <= r)[0]
            if len(ii) == 0 or len(jj) == 0 or len(kk) == 0:
                continue
            i0, i1, j0, j1, k0, k1 = ii[0], ii[-1] + 1, jj[0], jj[-1] + 1, kk[0], kk[-1] + 1
            d2 = ((xc[i0:i1, None, None] - shape.cx) ** 2 + (yc[None, j0:j1, None] - shape.cy) ** 2
                  + (zc[None, None, k0:k1] - shape.cz) ** 2)
            f = (d2 <= r * r).astype(np.float64)
        else:
            raise TypeError(f"unsupported shape {type(shape).__name__}")
        sub = epsr[i0:i1, j0:j1, k0:k1]
        hit = f > 0.0
        sub[hit] = (1.0 - f[hit]) * sub[hit] + f[hit] * mat.epsr
        if mat.model or mat.dispersive:
            li, lj, lk = np.nonzero(hit)
            lin = (li + i0).astype(np.int64) + nx * ((lj + j0) + ny * (lk + k0).astype(np.int64))
            (model_parts if mat.model else pole_parts).append((lin, f[hit]))
    inv_eps = 1.0 / epsr
    flat_inv = inv_eps.reshape(-1, order="F")

    def pack(parts):
        # unique sorted linear indices, keeping the largest fill of overlapping entries
        if not parts:
            return np.zeros(0, np.int64), np.zeros(0), np.zeros(0)
        lin = np.concatenate([p[0] for p in parts])
        fv = np.concatenate([p[1] for p in parts])
        keys, inv_idx = np.unique(lin, return_inverse=True)
        fills = np.zeros(len(keys))
        np.maximum.at(fills, inv_idx, fv)
        return keys, fills, flat_inv[keys]

    mc, mf, mi = pack(model_parts)
    pc, pf, pi = pack(pole_parts)
    return Geometry(epsr, inv_eps, mc, mf, mi, pc, pf, pi)
