"""Device-resident monitors: running DFT + net flux on planes, plane flux series, traces,
film reductions and their series monitors.

Mirrors the device path of src/sim/ProbeReadout.jl (`_probe_init!` :123-170,
`_record_probe_device!` :274-301) generalised from two x-normal planes to any set of up
to 8 planes on any axis, `record!(::FluxMonitor)` (src/sim/Monitors.jl:263-341) as an
on-device reduction, FilmAverage / SwitchedFraction / HotCellTrace (Monitors.jl:487-578)
with their records kept on the device, and switching_metrics (Monitors.jl:682-727).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, abi
from .state import DeviceBuffer, FDTDState


def _plane_dims(shape, axis: int):
    nx, ny, nz = shape
    return ((ny, nz), (nz, nx), (nx, ny))[axis]


def _cell_widths(grid, axis: int):
    return (grid.x, grid.y, grid.z)[axis].widths()


def plane_area(grid, axis: int) -> np.ndarray:
    """dA of every plane cell in plane order (n1, n2), Fortran-flattened
    (ProbeReadout.jl:139-141 for x; cyclic for y and z)."""
    a1, a2 = ((1, 2), (2, 0), (0, 1))[axis]
    w1, w2 = _cell_widths(grid, a1), _cell_widths(grid, a2)
    return np.outer(w1, w2).reshape(-1, order="F")


class DFTPlanes:
    """Per-cell running FP64 DFT of the tangential (Ea, Eb, Ha, Hb) and the integrated
    signed flux on each plane; `record(state)` once per step after step()."""

    def __init__(self, state: FDTDState, planes, omega_bins):
        if len(planes) > abi.MP_MAX_PLANES:
            raise ValueError(f"at most {abi.MP_MAX_PLANES} planes")
        self.state = state
        self.planes = [(int(a), int(i)) for a, i in planes]
        self.omega = np.asarray(omega_bins, dtype=np.float64)
        self.nb = len(self.omega)
        self.bufs = []
        arr = (abi.Plane * max(1, len(self.planes)))()
        for p, (axis, idx) in enumerate(self.planes):
            n1, n2 = _plane_dims(state.spec.shape, axis)
            npc = n1 * n2
            b = {k: DeviceBuffer.zeros(npc * self.nb * 16) for k in ("ea", "eb", "ha", "hb")}
            b["energy"] = DeviceBuffer.zeros(npc * 8)
            self.bufs.append(b)
            arr[p].axis, arr[p].index = axis, idx
            arr[p].dft_ea, arr[p].dft_eb = b["ea"].ptr, b["eb"].ptr
            arr[p].dft_ha, arr[p].dft_hb = b["ha"].ptr, b["hb"].ptr
            arr[p].energy = b["energy"].ptr
        self.arr = arr
        self.omega_dev = DeviceBuffer.from_array(self.omega)
        self.phase_dev = DeviceBuffer.zeros(16 * max(1, self.nb))

    def record(self, state: FDTDState = None) -> None:
        st = state or self.state
        _lib.call("probe_dft_accumulate", C.byref(st.yee), self.arr, len(self.planes), self.omega_dev.vp,
                  self.phase_dev.vp, self.nb, float(st.t), float(st.dt))

    def download(self, p: int) -> dict:
        axis, _ = self.planes[p]
        n1, n2 = _plane_dims(self.state.spec.shape, axis)
        out = {}
        for k in ("ea", "eb", "ha", "hb"):
            v = self.bufs[p][k].download(np.float64)
            out[k] = (v[0::2] + 1j * v[1::2]).reshape((n1, n2, self.nb), order="F")
        out["energy"] = self.bufs[p]["energy"].download(np.float64).reshape((n1, n2), order="F")
        return out

    def free(self):
        for b in self.bufs:
            for v in b.values():
                v.free()
        self.omega_dev.free()
        self.phase_dev.free()


class FluxSeries:
    """Plane-integrated Poynting flux per step, accumulated on the device."""

    def __init__(self, state: FDTDState, axis: int, index: int, capacity: int):
        self.state, self.axis, self.index = state, int(axis), int(index)
        self.area = DeviceBuffer.from_array(plane_area(state.spec.grid, self.axis))
        self.out = DeviceBuffer.zeros(8 * int(capacity))
        self.capacity = int(capacity)
        self.count = 0

    def due(self, n: int) -> bool:
        return True

    def record(self, state: FDTDState = None) -> None:
        st = state or self.state
        if self.count >= self.capacity:
            raise ValueError("flux series capacity exceeded")
        _lib.call("flux_plane", C.byref(st.yee), self.axis, self.index, self.area.vp, self.out.vp, self.count)
        self.count += 1

    def values(self) -> np.ndarray:
        return self.out.download(np.float64)[: self.count]


def film_reduce(state: FDTDState) -> dict:
    """Film means, switched fraction (m_TM_x < 0) and peak Te, reduced on device."""
    out = DeviceBuffer.zeros(80)
    _lib.call("film_reduce", C.byref(state.film), out.vp)
    v = out.download(np.float64)
    out.free()
    keys = ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z", "Te", "Tl", "switched", "Te_max")
    return dict(zip(keys, v.tolist()))


# ---- film averages, switched fraction, hot cell, switching diagnostics ----------------------

FILM_AVERAGE_KEYS = (
    "mag_time", "mx_TM", "mx_RE",
    "m_TM_x_reduced", "m_TM_y_reduced", "m_TM_z_reduced",
    "m_RE_x_reduced", "m_RE_y_reduced", "m_RE_z_reduced",
    "M_TM_x_Apm", "M_TM_y_Apm", "M_TM_z_Apm",
    "M_RE_x_Apm", "M_RE_y_Apm", "M_RE_z_Apm",
    "M_net_x_Apm", "M_net_y_Apm", "M_net_z_Apm",
    "Te_avg", "Tl_avg", "Ts_avg", "Te_K", "Tl_K", "Ts_TM_K", "Ts_RE_K",
    "mag_TM_norm_avg", "mag_RE_norm_avg",
)


def film_values_from_sums(sums, n: int, Ms_TM: float = 1.0, Ms_RE: float = 1.0) -> dict:
    """_film_state_values (Monitors.jl:522-556) from the 13 sums of mp_film_average:
    means are sum / n, NaN for an empty film."""
    s = np.asarray(sums, np.float64)
    mean = (lambda v: v / n) if n > 0 else (lambda v: float("nan"))
    mtx, mty, mtz, mrx, mry, mrz = (mean(s[i]) for i in range(6))
    Te, Tl, TsT, TsR = (mean(s[i]) for i in range(8, 12))
    MT = (Ms_TM * mtx, Ms_TM * mty, Ms_TM * mtz)
    MR = (Ms_RE * mrx, Ms_RE * mry, Ms_RE * mrz)
    return dict(mx_TM=mtx, mx_RE=mrx, m_TM_x_reduced=mtx, m_TM_y_reduced=mty, m_TM_z_reduced=mtz,
                m_RE_x_reduced=mrx, m_RE_y_reduced=mry, m_RE_z_reduced=mrz,
                M_TM_x_Apm=MT[0], M_TM_y_Apm=MT[1], M_TM_z_Apm=MT[2],
                M_RE_x_Apm=MR[0], M_RE_y_Apm=MR[1], M_RE_z_Apm=MR[2],
                M_net_x_Apm=MT[0] + MR[0], M_net_y_Apm=MT[1] + MR[1], M_net_z_Apm=MT[2] + MR[2],
                Te_avg=Te, Tl_avg=Tl, Ts_avg=0.5 * (TsT + TsR), Te_K=Te, Tl_K=Tl, Ts_TM_K=TsT, Ts_RE_K=TsR,
                mag_TM_norm_avg=mean(s[6]), mag_RE_norm_avg=mean(s[7]))


class _FilmSeries:
    """A film monitor whose records stay on the device (13 sums per record)."""

    def __init__(self, every: int = 1):
        self.every = int(every)
        self.t = []
        self.n_cells = None
        self._slots = None

    def due(self, n: int) -> bool:
        return self.every <= 1 or n % self.every == 0

    def record(self, state: FDTDState):
        from .probe import _Slots
        if self._slots is None:
            self._slots = _Slots(13)
        self.n_cells = 0 if state.film is None else int(state.film.n)
        if state.film is not None and self.n_cells > 0:
            _lib.call("film_average", C.byref(state.film), self._slots.next_ptr())
        else:
            self._slots.host.append(np.full((1, 13), np.nan))
        self.t.append(float(state.t))
        self._ms = (float(state.spec.gd["Ms_TM"]), float(state.spec.gd["Ms_RE"])) if state.spec.gd is not None \
            else (1.0, 1.0)
        return self

    def _sums(self) -> np.ndarray:
        return self._slots.values() if self._slots is not None else np.zeros((0, 13))

    def free(self):
        if self._slots is not None:
            self._slots.free()


class FilmAverage(_FilmSeries):
    """FilmAverage (Monitors.jl:113-116, 152-153, 558-570): film means every `every` steps."""

    def due(self, n: int) -> bool:
        return n % self.every == 0

    def data(self) -> dict:
        out = {k: [] for k in FILM_AVERAGE_KEYS}
        out["mag_time"] = list(self.t)
        ms = getattr(self, "_ms", (1.0, 1.0))
        for row in self._sums():
            v = film_values_from_sums(row, self.n_cells, *ms) if np.all(np.isfinite(row)) else \
                {k: float("nan") for k in FILM_AVERAGE_KEYS}
            for k in FILM_AVERAGE_KEYS[1:]:
                out[k].append(v[k])
        return {k: np.asarray(v, np.float64) for k, v in out.items()}


class SwitchedFraction(_FilmSeries):
    """SwitchedFraction (Monitors.jl:87-90, 137, 487-497): fraction of cells with m_TM_x < 0."""

    def data(self) -> dict:
        s = self._sums()
        vals = s[:, 12] / self.n_cells if self.n_cells else np.full(len(s), np.nan)
        return dict(t=np.asarray(self.t), values=np.asarray(vals, np.float64))


class HotCellTrace:
    """HotCellTrace (Monitors.jl:138-139, 499-512): Te, m_TM_x, m_RE_x of one film cell
    (0-based here), read back per record as the reference does."""

    def __init__(self, cell: int, every: int = 1):
        self.cell, self.every = int(cell), int(every)
        self.t, self.Te_K, self.mx_TM, self.mx_RE = [], [], [], []

    def due(self, n: int) -> bool:
        return self.every <= 1 or n % self.every == 0

    def record(self, state: FDTDState):
        self.t.append(float(state.t))
        f = state.film
        if f is None or not (0 <= self.cell < int(f.n)):
            for v in (self.Te_K, self.mx_TM, self.mx_RE):
                v.append(float("nan"))
            return self
        dt = np.dtype(state.spec.compute_dtype)
        out = np.empty(1, dt)
        for arr, ptr in ((self.Te_K, f.Te), (self.mx_TM, f.m_TM_x), (self.mx_RE, f.m_RE_x)):
            _lib.call("memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr + self.cell * dt.itemsize),
                      C.c_size_t(dt.itemsize))
            arr.append(float(out[0]))
        return self

    def data(self) -> dict:
        return dict(t=np.asarray(self.t), Te_K=np.asarray(self.Te_K), mx_TM=np.asarray(self.mx_TM),
                    mx_RE=np.asarray(self.mx_RE), cell=self.cell)

    def free(self):
        pass


def film_average(state: FDTDState) -> dict:
    """One FilmAverage record of the current film, reduced on the device."""
    n = 0 if state.film is None else int(state.film.n)
    if n == 0:
        return film_values_from_sums(np.zeros(13), 0)
    out = DeviceBuffer.zeros(8 * 13)
    _lib.call("film_average", C.byref(state.film), out.vp)
    s = out.download(np.float64)
    out.free()
    gd = state.spec.gd
    return film_values_from_sums(s, n, float(gd["Ms_TM"]), float(gd["Ms_RE"]))


def switching_from_sums(s, n: int, core_fraction: float = 0.5) -> dict:
    """switching_metrics (Monitors.jl:682-727) from the 13 counts and sums of
    mp_switching_metrics; hot_cell_index is 0-based."""
    nan = float("nan")
    if n <= 0:
        return dict(final_switch_fraction=nan, final_tm_reversed_fraction=nan, final_re_reversed_fraction=nan,
                    final_switched_cell_mean_m_TM_x=nan, final_switched_cell_mean_m_RE_x=nan,
                    final_unswitched_cell_mean_m_TM_x=nan, final_unswitched_cell_mean_m_RE_x=nan,
                    final_mixture_check_m_TM_x=nan, final_mixture_check_m_RE_x=nan,
                    core_Uabs_fraction_of_max=float(core_fraction), final_core_cell_count=0,
                    final_core_switch_fraction=nan, final_core_mean_m_TM_x=nan, final_core_mean_m_RE_x=nan,
                    hot_cell_index=0)
    s = np.asarray(s, np.float64)
    sw, un = int(s[0]), n - int(s[0])
    mm = lambda v, c: v / c if c > 0 else nan
    sw_frac = sw / n
    sw_tm, sw_re, un_tm, un_re = mm(s[3], sw), mm(s[4], sw), mm(s[5], un), mm(s[6], un)
    core = int(s[9])
    return dict(final_switch_fraction=sw_frac, final_tm_reversed_fraction=int(s[1]) / n,
                final_re_reversed_fraction=int(s[2]) / n,
                final_switched_cell_mean_m_TM_x=sw_tm, final_switched_cell_mean_m_RE_x=sw_re,
                final_unswitched_cell_mean_m_TM_x=un_tm, final_unswitched_cell_mean_m_RE_x=un_re,
                final_mixture_check_m_TM_x=sw_frac * sw_tm + (1.0 - sw_frac) * un_tm,
                final_mixture_check_m_RE_x=sw_frac * sw_re + (1.0 - sw_frac) * un_re,
                core_Uabs_fraction_of_max=float(core_fraction), final_core_cell_count=core,
                final_core_switch_fraction=(int(s[10]) / core if core > 0 else nan),
                final_core_mean_m_TM_x=mm(s[11], core), final_core_mean_m_RE_x=mm(s[12], core),
                hot_cell_index=int(s[8]))


def switching_metrics(state: FDTDState, core_fraction: float = 0.5) -> dict:
    """switching_metrics of the current film, with the counts and sums on the device."""
    n = 0 if state.film is None else int(state.film.n)
    if n == 0:
        return switching_from_sums(None, 0, core_fraction)
    out = DeviceBuffer.zeros(8 * 13)
    _lib.call("switching_metrics", C.byref(state.film), float(core_fraction), out.vp)
    s = out.download(np.float64)
    out.free()
    return switching_from_sums(s, n, core_fraction)
