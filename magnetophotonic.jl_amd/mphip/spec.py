"""Host-side construction of a coupled-step state (the FDTDState constructor).

Restates the host half of src/fdtd/Solver.jl:55-193: compute-type rounding of the
model parameters (convert_compute_model), absolute 1/(eps0 eps_r) volumes, CT copies of
the 1-D metric arrays, CPML coefficients, per-component ADE lists with 1/(eps0 eps)
screening factors, position maps, pole fits, initial thermal / magnetisation /
absorption state and the m_eq LUT.  Everything here is a host numpy array; `FDTDState`
(state.py) uploads it to the device.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import dispersion, models
from .cpml import CPMLSpec, build_cpml
from .geometry import Geometry
from .grid import FDTDParams, Grid3D

DTYPE_NAMES = {"f32": np.float32, "f64": np.float64}


def _np_dtype(spec) -> np.dtype:
    if isinstance(spec, str):
        if spec not in DTYPE_NAMES:
            raise ValueError("compute_precision must be 'f32' or 'f64'")
        return np.dtype(DTYPE_NAMES[spec])
    d = np.dtype(spec)
    if d not in (np.float32, np.float64):
        raise ValueError("compute_precision must be float32 or float64")
    return d


def position_map(cells: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """active_axis_position_map (src/geometry/ReferenceRaster.jl:218-224), 0-based with
    -1 for a cell absent from `idx`."""
    cells = np.asarray(cells, dtype=np.int64)
    idx = np.asarray(idx, dtype=np.int64)
    if len(idx) == 0 or len(cells) == 0:
        return np.full(len(cells), -1, np.int32)
    order = np.argsort(idx, kind="stable")
    sidx = idx[order]
    p = np.clip(np.searchsorted(sidx, cells), 0, len(sidx) - 1)
    return np.where(sidx[p] == cells, order[p], -1).astype(np.int32)


@dataclass
class ADEList:
    idx: np.ndarray    # int32, 0-based linear cell index
    fill: np.ndarray   # CT
    inv: np.ndarray    # CT, 1/(eps0 eps) of the cell


@dataclass
class StateSpec:
    grid: Grid3D
    dt: float
    params: FDTDParams
    field_dtype: np.dtype
    compute_dtype: np.dtype
    psi_dtype: str                  # 'f32' | 'f64' | 'bf16' | 'f16'
    psi_scale: float
    inv_eps: tuple                  # 3 flat volumes (field dtype), absolute 1/(eps0 eps_r)
    inv_d_cell: tuple               # 3 CT vectors
    inv_d_dual: tuple
    cpml: CPMLSpec
    source: Any
    ade: Optional[tuple]            # (ADEList x, y, z) or None
    diag_poles: Optional[np.ndarray]  # npoles*3 CT
    ade_state_dtype: str
    ade_state_scale: float
    mo: Optional[ADEList]
    mo_pos: Optional[np.ndarray]
    mo_poles: Optional[np.ndarray]
    mo_pair: tuple                  # ('y', 'z') in the reference
    film_cells: np.ndarray          # int32
    pos: tuple                      # pos_x, pos_y, pos_z (int32, -1 = none)
    film_init: dict                 # Te, Tl, Ts_TM, Ts_RE, m_TM_x ... (CT arrays)
    gd: Optional[models.FerrimagnetParameters]   # CT-rounded parameters
    lut: Optional[tuple]            # (tm CT, re CT, T_min, inv_dT, n)
    multiphysics_every: int
    subcycles: int
    brillouin_iters: int
    multiphysics_enabled: bool
    absorption_model: str
    has_film: bool
    extras: dict = field(default_factory=dict)
    epsr: Optional[np.ndarray] = None   # (nx, ny, nz) CT relative permittivity, the state's epsr (Solver.jl:99)
    film_fill: Optional[np.ndarray] = None   # CT fill fraction of each film cell (region.material_fill, Solver.jl:148)

    @property
    def shape(self):
        return self.grid.shape

    @property
    def n_film(self) -> int:
        return int(len(self.film_cells))

    def pabs_factor(self) -> float:
        """0.5 * omega_pump * eps0 * eps_imag_pump (Kernels.jl:631)."""
        if self.gd is None:
            return 0.0
        return 0.5 * float(self.gd["omega_pump"]) * float(self.params.eps0) * float(self.gd["eps_imag_pump"])

    def has_ade_work(self) -> bool:
        """Kernels.jl:632-634."""
        return (self.absorption_model == "ade_work" and self.ade is not None
                and self.diag_poles is not None and len(self.ade[0].idx) * (len(self.diag_poles) // 3) > 0)


def build_state_spec(grid: Grid3D, geo: Geometry, dt: float, params: FDTDParams = FDTDParams(),
                     compute_precision="f32", field_precision="f32", source=None, n_pml=0,
                     cpml_kwargs=None, model: Optional[models.FerrimagnetParameters] = None,
                     enable_dispersion: Optional[bool] = None, enable_magneto_optic: bool = False,
                     diag_poles=None, mo_poles=None, multiphysics_every: int = 1, subcycles: int = 0,
                     brillouin_iters: int = 2, enable_multiphysics: bool = True,
                     absorption_model: str = "cycle_average", seed_tilt: float = 1e-4,
                     psi_precision: Optional[str] = None, psi_scale: float = 1.0,
                     ade_state_precision: Optional[str] = None, ade_state_scale: float = 1.0,
                     mo_pair=("y", "z"), dispersive_cells: str = "model") -> StateSpec:
    """FDTDState(grid, geo; ...) host construction (Solver.jl:55-193).

    `model` is the Float64 FerrimagnetParameters (None = NullModel).  `dispersive_cells`
    = "model" takes the ADE lists from the film cells, as the reference does; "poles"
    takes them from the dispersive-material cells and runs no film physics (the
    EM-only dispersive-sphere workload, which the reference state cannot express).
    """
    CT = _np_dtype(compute_precision)
    FT = _np_dtype(field_precision)
    cpml_kwargs = dict(cpml_kwargs or {})
    psi = psi_precision or ("f32" if CT == np.float32 else "f64")
    if psi not in ("f32", "f64", "bf16", "f16"):
        raise ValueError("psi precision must be f32, f64, bf16 or f16")
    gd = model.converted(CT) if model is not None else None       # convert_compute_model
    dtf = float(dt)
    cpml = build_cpml(grid, n_pml, dtf, params, dtype=CT, **cpml_kwargs)
    inv_eps0 = 1.0 / float(params.eps0)
    flat = geo.inv_eps.reshape(-1, order="F")
    inv_vol = (flat * inv_eps0).astype(FT)
    inv_eps = (inv_vol, inv_vol.copy(), inv_vol.copy())
    idc = tuple(a.inv_d_cell.astype(CT) for a in (grid.x, grid.y, grid.z))
    idd = tuple(a.inv_d_dual.astype(CT) for a in (grid.x, grid.y, grid.z))

    use_poles = dispersive_cells == "poles"
    if use_poles:
        cells = geo.pole_cells.astype(np.int64)
        fill = geo.pole_fill.astype(np.float64)
        inv_host = geo.pole_inv_eps.astype(np.float64) / float(params.eps0)
    else:
        cells = geo.material_cells.astype(np.int64)
        fill = geo.material_fill.astype(np.float64)
        inv_host = geo.material_inv_eps.astype(np.float64) / float(params.eps0)
    nmat = len(cells)
    has_model = gd is not None and nmat > 0
    if enable_dispersion is None:
        enable_dispersion = gd is not None

    if cells.size and cells.max() >= 2 ** 31:
        raise ValueError("cell index exceeds int32")
    idx32 = cells.astype(np.int32)
    lst = ADEList(idx32, fill.astype(CT), inv_host.astype(CT))
    ade = None
    dpoles = None
    if (has_model or (use_poles and nmat > 0)) and enable_dispersion:
        if diag_poles is None:
            if gd is None:
                raise ValueError("dispersive cells need diag_poles or a model")
            diag_poles = dispersion.build_pump_poles(dtf, params.eps0, gd)
        dpoles = dispersion.poles_array(diag_poles, CT)
        ade = (lst, ADEList(idx32.copy(), lst.fill.copy(), lst.inv.copy()),
               ADEList(idx32.copy(), lst.fill.copy(), lst.inv.copy()))

    has_film = has_model and not use_poles
    film_cells = idx32 if has_film else np.zeros(0, np.int32)
    pos = tuple(position_map(film_cells, a.idx) if (ade is not None and has_film) else
                np.full(len(film_cells), -1, np.int32) for a in (ade or (lst, lst, lst)))
    film_init = {}
    lut = None
    mo = mo_pos = mpoles = None
    if has_film:
        tm, re_, T_min, inv_dT, nT = models.build_m_eq_lut(gd)
        lut = (tm.astype(CT), re_.astype(CT), float(CT.type(T_min)), float(CT.type(inv_dT)), nT)
        T0 = CT.type(gd["T0"])
        mtm0 = CT.type(models.lookup_m_eq_lut(tm, gd["T0"], T_min, inv_dT, nT))
        mre0 = CT.type(models.lookup_m_eq_lut(re_, gd["T0"], T_min, inv_dT, nT))
        st = CT.type(seed_tilt)
        n = len(film_cells)
        film_init = dict(Te=np.full(n, T0, CT), Tl=np.full(n, T0, CT), Ts_TM=np.full(n, T0, CT),
                         Ts_RE=np.full(n, T0, CT), m_TM_x=np.full(n, mtm0, CT),
                         m_TM_y=np.zeros(n, CT), m_TM_z=np.full(n, st, CT),
                         m_RE_x=np.full(n, mre0, CT), m_RE_y=np.zeros(n, CT),
                         m_RE_z=np.full(n, st, CT))
        if enable_magneto_optic:
            if mo_poles is None:
                mo_poles = dispersion.build_probe_mo_poles(dtf, params.eps0, gd)
            mpoles = dispersion.poles_array(mo_poles, CT)
            comp_b = {"x": 0, "y": 1, "z": 2}[mo_pair[1]]
            base = ade[comp_b] if ade is not None else lst
            mo = ADEList(base.idx.copy(), base.fill.copy(), base.inv.copy())
            mo_pos = position_map(mo.idx, film_cells)
    subs = subcycles if subcycles > 0 else max(1, multiphysics_every)
    ade_sd = ade_state_precision or ("f32" if CT == np.float32 else "f64")
    return StateSpec(grid=grid, dt=dtf, params=params, field_dtype=FT, compute_dtype=CT,
                     psi_dtype=psi, psi_scale=float(psi_scale), inv_eps=inv_eps, inv_d_cell=idc,
                     inv_d_dual=idd, cpml=cpml, source=source, ade=ade, diag_poles=dpoles,
                     ade_state_dtype=ade_sd, ade_state_scale=float(ade_state_scale), mo=mo,
                     mo_pos=mo_pos, mo_poles=mpoles, mo_pair=tuple(mo_pair), film_cells=film_cells,
                     pos=pos, film_init=film_init, gd=gd, lut=lut,
                     multiphysics_every=int(multiphysics_every), subcycles=int(subs),
                     brillouin_iters=int(brillouin_iters), multiphysics_enabled=bool(enable_multiphysics),
                     absorption_model=absorption_model, has_film=has_film,
                     epsr=np.asarray(geo.epsr, order="F").astype(CT, order="F"),
                     film_fill=(lst.fill.copy() if has_film else np.zeros(0, CT)))


def _small_codes(a: np.ndarray, limit: int):
    """Sorted distinct values of `a` and each element's position among them, or None when
    there are more than `limit` (a subsample proposes candidates; every element is then
    checked against them, so the result is exact)."""
    flat = np.ascontiguousarray(a).reshape(-1)
    cand = np.unique(flat[:: max(1, flat.size // 65536)])
    for _ in range(4):
        if cand.size > limit:
            return None
        pos = np.minimum(np.searchsorted(cand, flat), cand.size - 1)
        miss = cand[pos] != flat
        if not miss.any():
            return cand, pos
        cand = np.union1d(cand, np.unique(flat[miss]))
    return None


def eps_index(inv_eps, limit: int = 256):
    """Compact form of the three inv_eps volumes for the FP32 E update: a uint8 index
    per cell and a (n, 4) FP32 table of the distinct (x, y, z) triples, or None when
    there are more than `limit` distinct triples.  The table holds the volumes' own
    values, so the update that reads it is unchanged."""
    if any(np.asarray(a).dtype != np.float32 for a in inv_eps):
        return None
    codes = [_small_codes(a, limit) for a in inv_eps]
    if any(c is None for c in codes):
        return None
    (vx, px), (vy, py), (vz, pz) = codes
    key = (px.astype(np.int64) * vy.size + py) * vz.size + pz
    kc = _small_codes(key, limit)
    if kc is None:
        return None
    keys, index = kc
    table = np.zeros((keys.size, 4), np.float32)
    table[:, 0] = vx[keys // (vy.size * vz.size)]
    table[:, 1] = vy[(keys // vz.size) % vy.size]
    table[:, 2] = vz[keys % vz.size]
    return index.astype(np.uint8), table
