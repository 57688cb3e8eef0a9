"""Host-side field diagnostics (src/fdtd/Fields.jl:89-100)."""
from __future__ import annotations

import numpy as np

from .grid import FDTDParams, Grid3D


def field_energy(fields: dict, grid: Grid3D, p: FDTDParams = FDTDParams()) -> float:
    """field_energy(::FieldState, ::Grid3D): sum over cells of
    0.5 (eps0 |E|^2 + mu0 |H|^2) dx dy dz, with `fields` mapping ex..hz to (nx, ny, nz)."""
    vol = (grid.x.widths()[:, None, None] * grid.y.widths()[None, :, None] * grid.z.widths()[None, None, :])
    e2 = sum(np.asarray(fields[k], np.float64) ** 2 for k in ("ex", "ey", "ez"))
    h2 = sum(np.asarray(fields[k], np.float64) ** 2 for k in ("hx", "hy", "hz"))
    return float(np.sum(0.5 * (p.eps0 * e2 + p.mu0 * h2) * vol))
