"""Device-resident monitors: running DFT + net flux on planes, plane flux series, traces
and film reductions.

Mirrors the device path of src/sim/ProbeReadout.jl (`_probe_init!` :123-170,
`_record_probe_device!` :274-301) generalised from two x-normal planes to any set of up
to 8 planes on any axis, `record!(::FluxMonitor)` (src/sim/Monitors.jl:263-341) as an
on-device reduction, and FilmAverage / SwitchedFraction (Monitors.jl:487-578).
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
