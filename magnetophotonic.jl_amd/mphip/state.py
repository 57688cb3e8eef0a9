"""FDTDState on the MI355X: device-resident coupled-step state and the step sequence.

Mirrors src/fdtd/Solver.jl: the data model (:9-53), `step!`/`_step_typed!` (:277-369),
the multiphysics window (:373-390), `relax_step!` (:548-555), `run!` (:557-563),
`magnetization_snapshot` / `apply_magnetization!` (:448-483), `configure_probe_mode!`
(:416-443) and `free_state!` (:521-546).  Every per-step operation is a libmphip call on
one stream; the host never touches field data inside the loop.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import _lib, abi, dispersion
from .sources import (AXIS_CODE, COMPONENT_CODE, CarrierGaussianPulse, ContinuousSource, GaussianPulse,
                      ModeSource, PlaneSource, SoftSource, axis_weights, source_index)
from .spec import StateSpec, eps_index

_DT_CODE = {np.dtype(np.float32): abi.MP_F32, np.dtype(np.float64): abi.MP_F64}
_PSI_CODE = {"f32": abi.MP_F32, "f64": abi.MP_F64, "bf16": abi.MP_BF16, "f16": abi.MP_F16}
_PSI_SIZE = {"f32": 4, "f64": 8, "bf16": 2, "f16": 2}


class DeviceBuffer:
    """A device allocation owned by Python (finalised by `free`)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _lib.call("malloc", C.byref(p), C.c_size_t(max(self.nbytes, 0)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(a.nbytes)
        if a.nbytes:
            _lib.call("memcpy_h2d", C.c_void_p(b.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))
        return b

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        b = cls(nbytes)
        if nbytes:
            _lib.call("memset", C.c_void_p(b.ptr), 0, C.c_size_t(nbytes))
        return b

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {a.nbytes} vs {self.nbytes}")
        if a.nbytes:
            _lib.call("memcpy_h2d", C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))

    def download(self, dtype, count: Optional[int] = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dtype)
        if out.nbytes:
            _lib.call("memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.call("free", C.c_void_p(self.ptr))
            self.ptr = 0

    @property
    def vp(self):
        return C.c_void_p(self.ptr) if self.ptr else None


FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")
FILM_NAMES = ("Te", "Tl", "Ts_TM", "Ts_RE", "m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z")


YEE_FORMS = {"auto": abi.MP_YEE_AUTO, "whole": abi.MP_YEE_WHOLE, "scalar": abi.MP_YEE_SCALAR}


class FDTDState:
    """Device state of one coupled 3-D simulation; `step()` advances one EM step.

    use_eps_index: give the FP32 E update the 1-byte material index where the volumes hold
    at most 256 distinct triples (False: it reads the three inv_eps volumes, as the Julia
    extension's launcher does).  yee_form: "auto", "whole" or "scalar" kernel family of
    the FP32 Yee updates.  film_serial: run the film update in step order on the launch
    stream instead of beside the next step's Yee updates.  pingpong: allocate a second set
    of the six E and H volumes, with which `run` advances FP32 states in pairs of ping-pong
    steps whose interior H and E updates are one fused launch (where the grid and source
    allow it; the library falls back to in-place steps otherwise).  All of these give the
    same values.  Measured on the 512 x 512 x 256 pump step the ping-pong form is slower
    (1.78 against 1.52 ms), so it is off by default."""

    def __init__(self, spec: StateSpec, device: int = 0, use_eps_index: bool = True, yee_form: str = "auto",
                 film_serial: bool = False, pingpong: bool = False):
        _lib.init(device)
        self.spec = spec
        self._st = abi.State()
        self._film_serial = film_serial
        self._dft = None
        self.n = 0
        self.t = 0.0
        self.last_mp_n = 0
        self.multiphysics_enabled = spec.multiphysics_enabled
        self.dt = spec.dt
        self._bufs = []
        nx, ny, nz = spec.shape
        ncell = nx * ny * nz
        FT, CT = spec.field_dtype, spec.compute_dtype
        self._fcode, self._ccode = _DT_CODE[FT], _DT_CODE[CT]

        def dev_zeros(nbytes):
            b = DeviceBuffer.zeros(nbytes)
            self._bufs.append(b)
            return b

        def dev_arr(a):
            b = DeviceBuffer.from_array(a)
            self._bufs.append(b)
            return b

        # Yee volumes and metric arrays
        self.f = {k: dev_zeros(ncell * FT.itemsize) for k in FIELD_NAMES}
        self.inv_eps = [dev_arr(a) for a in spec.inv_eps]
        self.idc = [dev_arr(a) for a in spec.inv_d_cell]
        self.idd = [dev_arr(a) for a in spec.inv_d_dual]
        y = abi.Yee()
        y.nx, y.ny, y.nz = nx, ny, nz
        y.field_dtype, y.compute_dtype = self._fcode, self._ccode
        for k in FIELD_NAMES:
            setattr(y, k, self.f[k].ptr)
        y.inv_eps_x, y.inv_eps_y, y.inv_eps_z = (b.ptr for b in self.inv_eps)
        y.inv_d_cell_x, y.inv_d_cell_y, y.inv_d_cell_z = (b.ptr for b in self.idc)
        y.inv_d_dual_x, y.inv_d_dual_y, y.inv_d_dual_z = (b.ptr for b in self.idd)
        y.kernel_form = YEE_FORMS[yee_form]
        ei = eps_index(spec.inv_eps) if use_eps_index and FT == np.float32 and CT == np.float32 else None
        if ei is not None:
            self._eps_index, self._eps_table = dev_arr(ei[0]), dev_arr(ei[1])
            y.eps_index, y.eps_table, y.n_eps = self._eps_index.ptr, self._eps_table.ptr, ei[1].shape[0]
        self.yee = y
        self.yee_alt = None
        if pingpong and FT == np.float32 and CT == np.float32 and yee_form == "auto":
            self.yee_alt = [dev_zeros(ncell * FT.itemsize) for _ in range(6)]

        # CPML
        cs = spec.cpml
        cp = abi.Cpml()
        cp.npml_x, cp.npml_y, cp.npml_z = cs.widths
        cp.psi_dtype = _PSI_CODE[spec.psi_dtype]
        cp.psi_scale = spec.psi_scale
        self._cpml_coef = []
        for d, ax in enumerate(cs.axes):
            bk, ba, bb = dev_arr(ax.inv_kappa), dev_arr(ax.a), dev_arr(ax.b)
            cp.inv_kappa[d], cp.a[d], cp.b[d] = bk.ptr, ba.ptr, bb.ptr
        from .cpml import PSI_AXIS
        esz = _PSI_SIZE[spec.psi_dtype]
        self.psi_h, self.psi_d = [], []
        for s in range(12):
            nb = cs.slab_cells(PSI_AXIS[s]) * esz
            bh, bd = dev_zeros(nb), dev_zeros(nb)
            self.psi_h.append(bh)
            self.psi_d.append(bd)
            cp.psi_h[s] = bh.ptr or None
            cp.psi_d[s] = bd.ptr or None
        self.cpml = cp

        # source
        self._mode_profile = None
        src = spec.source
        if isinstance(src, ModeSource):
            if src.axis != "x":
                raise ValueError("ModeSource injection supports axis='x'")  # Kernels.jl:581
            if src.component not in ("Ey", "Ez"):
                raise ValueError("ModeSource electric component must be Ey or Ez")
            prof = np.asarray(src.profile, dtype=np.float64).reshape(-1, order="F").astype(CT)
            self._mode_profile = dev_arr(prof)
            self._mode_i = source_index(spec.grid.x, src.position)
            eta0 = math.sqrt(spec.params.mu0 / spec.params.eps0)
            self._mode_hcorr = src.neff / eta0

        # ADE lists
        self.ade = None
        self.ade_bufs = None
        if spec.ade is not None:
            np_ = len(spec.diag_poles) // 3
            self.poles_dev = dev_arr(spec.diag_poles)
            sd = spec.ade_state_dtype
            ssz = _PSI_SIZE[sd]
            self.ade = []
            self.ade_bufs = []
            for lst in spec.ade:
                n = len(lst.idx)
                bufs = dict(P=dev_zeros(n * np_ * ssz), J=dev_zeros(n * np_ * ssz),
                            E_old=dev_zeros(n * CT.itemsize), idx=dev_arr(lst.idx),
                            fill=dev_arr(lst.fill), inv=dev_arr(lst.inv))
                a = abi.Ade()
                a.n, a.npoles, a.dtype = n, np_, self._ccode
                a.state_dtype, a.state_scale = _PSI_CODE[sd], spec.ade_state_scale
                a.P, a.J, a.E_old = bufs["P"].ptr, bufs["J"].ptr, bufs["E_old"].ptr
                a.idx, a.fill, a.inv_eps = bufs["idx"].ptr, bufs["fill"].ptr, bufs["inv"].ptr
                a.poles = self.poles_dev.ptr
                self.ade.append(a)
                self.ade_bufs.append(bufs)

        # film state
        self.film = None
        if spec.has_film:
            n = spec.n_film
            self.film_bufs = {k: dev_arr(spec.film_init[k]) for k in FILM_NAMES}
            self.film_bufs["cells"] = dev_arr(spec.film_cells)
            for c, pos in zip("xyz", spec.pos):
                self.film_bufs["pos_" + c] = dev_arr(pos)
            for k in ("U_abs", "window_sum", "peak"):
                self.film_bufs[k] = dev_zeros(n * 8)
            fm = abi.Film()
            fm.n, fm.dtype = n, self._ccode
            for k, b in self.film_bufs.items():
                setattr(fm, k, b.ptr)
            self.film = fm
            self.gd_arr = np.ascontiguousarray(spec.gd.as_array(np.float64))
            tm, re_, T_min, inv_dT, nT = spec.lut
            self.lut_tm, self.lut_re = dev_arr(tm), dev_arr(re_)
            lut = abi.Lut()
            lut.tm, lut.re, lut.T_min, lut.inv_dT, lut.n, lut.dtype = (
                self.lut_tm.ptr, self.lut_re.ptr, T_min, inv_dT, nT, self._ccode)
            self.lut = lut
            self.Q_TM = float(spec.gd["Q_voigt_TM"])
            self.Q_RE = float(spec.gd["Q_voigt_RE"])

        # magneto-optic gyration
        self.mo = None
        if spec.mo is not None:
            self._build_mo(spec.mo_poles)
        self._build_state()

    # ---- construction helpers --------------------------------------------------------------
    def _build_mo(self, mpoles: np.ndarray):
        spec = self.spec
        CT = spec.compute_dtype
        n = len(spec.mo.idx)
        np_ = len(mpoles) // 3
        self.mo_poles_dev = DeviceBuffer.from_array(mpoles)
        self._bufs.append(self.mo_poles_dev)
        mb = {k: DeviceBuffer.zeros(n * np_ * CT.itemsize) for k in ("Pa", "Ja", "Pb", "Jb")}
        mb.update(Ea_old=DeviceBuffer.zeros(n * CT.itemsize), Eb_old=DeviceBuffer.zeros(n * CT.itemsize),
                  idx=DeviceBuffer.from_array(spec.mo.idx), fill=DeviceBuffer.from_array(spec.mo.fill),
                  mpos=DeviceBuffer.from_array(spec.mo_pos))
        self._bufs.extend(mb.values())
        m = abi.Mo()
        m.n, m.npoles, m.dtype = n, np_, self._ccode
        m.Pa_from_b, m.Ja_from_b, m.Pb_from_a, m.Jb_from_a = mb["Pa"].ptr, mb["Ja"].ptr, mb["Pb"].ptr, mb["Jb"].ptr
        m.E_old_a, m.E_old_b = mb["Ea_old"].ptr, mb["Eb_old"].ptr
        m.idx, m.fill, m.mpos, m.poles = mb["idx"].ptr, mb["fill"].ptr, mb["mpos"].ptr, self.mo_poles_dev.ptr
        self.mo = m
        self.mo_bufs = mb
        a, b = spec.mo_pair
        self._mo_ea, self._mo_eb = self.f["e" + a], self.f["e" + b]
        ci = {"x": 0, "y": 1, "z": 2}
        self._mo_iea, self._mo_ieb = self.inv_eps[ci[a]], self.inv_eps[ci[b]]

    # ---- the step (Solver.jl:283-383) --------------------------------------------------------
    # n, t, last_mp_n and the multiphysics flag live in the mp_state the C driver advances.
    @property
    def n(self) -> int:
        return int(self._st.n)

    @n.setter
    def n(self, v):
        self._st.n = int(v)

    @property
    def t(self) -> float:
        return float(self._st.t)

    @t.setter
    def t(self, v):
        self._st.t = float(v)

    @property
    def last_mp_n(self) -> int:
        return int(self._st.last_mp_n)

    @last_mp_n.setter
    def last_mp_n(self, v):
        self._st.last_mp_n = int(v)

    @property
    def multiphysics_enabled(self) -> bool:
        return bool(self._st.multiphysics_enabled)

    @multiphysics_enabled.setter
    def multiphysics_enabled(self, v):
        self._st.multiphysics_enabled = 1 if v else 0

    def _source_struct(self) -> abi.Source:
        src = self.spec.source
        s = abi.Source()
        s.kind = abi.MP_SRC_NONE
        if src is None:
            return s
        p = src.pulse
        if isinstance(p, GaussianPulse):
            s.pulse.kind, s.pulse.amplitude, s.pulse.t0 = abi.MP_PULSE_GAUSSIAN, p.amplitude, p.t0
            s.pulse.tau, s.pulse.omega, s.pulse.phase = p.tau, p.omega, p.phase
        elif isinstance(p, CarrierGaussianPulse):
            s.pulse.kind, s.pulse.amplitude, s.pulse.t0 = abi.MP_PULSE_CARRIER_GAUSSIAN, p.amplitude, p.t0
            s.pulse.tau, s.pulse.omega = p.sigma, p.omega
        elif isinstance(p, ContinuousSource):
            s.pulse.kind, s.pulse.amplitude, s.pulse.omega, s.pulse.phase = (
                abi.MP_PULSE_CONTINUOUS, p.amplitude, p.omega, p.phase)
        else:
            raise TypeError(f"unsupported pulse {type(p).__name__}")
        s.component = COMPONENT_CODE[src.component]
        if isinstance(src, SoftSource):
            s.kind = abi.MP_SRC_SOFT
            s.i, s.j, s.k = (int(v) for v in src.index)
        elif isinstance(src, ModeSource):
            s.kind = abi.MP_SRC_MODE_X
            s.i = self._mode_i
            s.profile = self._mode_profile.ptr
            s.hcorr = self._mode_hcorr
        elif isinstance(src, PlaneSource):
            g = self.spec.grid
            w = axis_weights({"x": g.x, "y": g.y, "z": g.z}[src.axis], src.position)
            s.kind = abi.MP_SRC_PLANE
            s.axis = AXIS_CODE[src.axis]
            s.i0, s.w0 = w[0]
            s.i1, s.w1 = w[1] if len(w) > 1 else (-1, 0.0)
        else:
            raise TypeError(f"unsupported source {type(src).__name__}")
        return s

    def _build_state(self):
        """Fill the mp_state the C step driver consumes (everything one step touches)."""
        spec = self.spec
        st = self._st
        st.yee = C.pointer(self.yee)
        st.cpml = C.pointer(self.cpml)
        for c in range(3):
            st.ade[c] = C.pointer(self.ade[c]) if self.ade is not None else None
        if self.mo is not None and self.film is not None:
            st.mo = C.pointer(self.mo)
            st.mo_ea, st.mo_eb = self._mo_ea.ptr, self._mo_eb.ptr
            st.mo_iea, st.mo_ieb = self._mo_iea.ptr, self._mo_ieb.ptr
        else:
            st.mo = None
        if self.film is not None:
            st.film = C.pointer(self.film)
            st.params = self.gd_arr.ctypes.data
            st.lut = C.pointer(self.lut)
            st.Q_TM, st.Q_RE = self.Q_TM, self.Q_RE
        st.source = self._source_struct()
        st.dt, st.mu0, st.pabs_factor = self.dt, spec.params.mu0, spec.pabs_factor()
        st.multiphysics_every, st.subcycles = spec.multiphysics_every, spec.subcycles
        st.brillouin_iters = spec.brillouin_iters
        st.has_ade_work = 1 if spec.has_ade_work() else 0
        st.film_lists_coincide = 1 if self._lists_coincide() else 0
        st.film_serial = 1 if self._film_serial else 0
        for c in range(6):
            st.yee_alt[c] = self.yee_alt[c].ptr if self.yee_alt is not None else None

    def _lists_coincide(self) -> bool:
        """Every ADE list and the MO list is the film list (entry n is film cell n) and the
        ADE state is stored in the compute type, or in BF16/F16 with FP32 compute: the step
        may run its list work as one mp_film_epilogue pass (the lists FDTDState builds from
        the material cells, Solver.jl:114-159)."""
        spec = self.spec
        if self.film is None or spec.ade is None or spec.n_film == 0:
            return False
        ct = _DT_CODE[np.dtype(spec.compute_dtype)]
        sd = _PSI_CODE[spec.ade_state_dtype]
        if sd != ct and not (ct == abi.MP_F32 and sd in (abi.MP_BF16, abi.MP_F16)):
            return False
        ident = np.arange(spec.n_film, dtype=np.int64)
        for c in range(3):
            if not np.array_equal(spec.ade[c].idx, spec.film_cells):
                return False
            if not np.array_equal(spec.pos[c], ident):
                return False
        if spec.mo is not None and spec.mo_pos is not None:
            if not np.array_equal(spec.mo.idx, spec.film_cells) or not np.array_equal(spec.mo_pos, ident):
                return False
        return True

    @property
    def fused_epilogue(self) -> bool:
        """Whether the step runs its film list work as one mp_film_epilogue pass."""
        return bool(self._st.film_lists_coincide)

    @fused_epilogue.setter
    def fused_epilogue(self, on: bool) -> None:
        if on and not self._lists_coincide():
            raise ValueError("the ADE and MO lists are not the film list; the fused pass does not apply")
        self._st.film_lists_coincide = 1 if on else 0

    def set_source(self, src) -> None:
        """Replace the state's source (the phase-source swap of run_phase!,
        src/sim/Experiment.jl:254-290)."""
        self.spec.source = src
        if isinstance(src, ModeSource):
            if src.axis != "x" or src.component not in ("Ey", "Ez"):
                raise ValueError("ModeSource injection supports axis='x' with component Ey or Ez")
            prof = np.asarray(src.profile, dtype=np.float64).reshape(-1, order="F").astype(self.spec.compute_dtype)
            self._mode_profile = DeviceBuffer.from_array(prof)
            self._bufs.append(self._mode_profile)
            self._mode_i = source_index(self.spec.grid.x, src.position)
            self._mode_hcorr = src.neff / math.sqrt(self.spec.params.mu0 / self.spec.params.eps0)
        self._st.source = self._source_struct()

    def attach_dft(self, mon) -> None:
        """Record `mon` (monitors.DFTPlanes) inside every step of run()/step()."""
        st = self._st
        st.planes = C.cast(mon.arr, C.POINTER(abi.Plane))
        st.nplanes, st.nb = len(mon.planes), mon.nb
        st.omega_bins, st.phase = mon.omega_dev.ptr, mon.phase_dev.ptr
        st.record_dft = 1
        self._dft = mon

    def detach_dft(self) -> None:
        self._st.record_dft = 0
        self._st.nplanes = 0
        self._dft = None

    @property
    def film_serial(self) -> bool:
        """The film update runs on the main stream after the step's launches instead of on
        the side stream beside the next step's Yee updates (same results)."""
        return bool(self._st.film_serial)

    @film_serial.setter
    def film_serial(self, on: bool) -> None:
        self._film_serial = bool(on)
        self._st.film_serial = 1 if on else 0

    def step(self) -> "FDTDState":
        """One EM step (step!, Solver.jl:277-369)."""
        return self.run(1)

    def run(self, steps: int, callback=None):
        """run! (Solver.jl:557-563): the whole step sequence in one C call per step batch."""
        if callback is None:
            _lib.call("run_steps", C.byref(self._st), int(steps))
        else:
            for _ in range(int(steps)):
                _lib.call("run_steps", C.byref(self._st), 1)
                callback(self)
        return self

    def attach_surrogate(self, sur, tol: float = 1e-4) -> None:
        """Film updates by `sur` (surrogate.Surrogate) instead of the plain integrator, with
        a plain fallback wherever the network misses one plain substep by more than tol."""
        self._sur = sur
        self._st.surrogate = C.pointer(sur.mlp)
        self._st.surrogate_tol = float(tol)
        self._st.surrogate_fallbacks = sur.fallbacks.ptr

    def detach_surrogate(self) -> None:
        self._sur = None
        self._st.surrogate = None
        self._st.surrogate_fallbacks = None

    def _film_update(self, dt: float, subcycles: int, has_pabs: int) -> None:
        sur = getattr(self, "_sur", None)
        gd = self.gd_arr.ctypes.data_as(C.c_void_p)
        if sur is None:
            _lib.call("multiphysics_step", C.byref(self.film), gd, C.byref(self.lut), float(dt), int(subcycles),
                      self.spec.brillouin_iters, int(has_pabs))
        else:
            _lib.call("surrogate_step", C.byref(self.film), gd, C.byref(self.lut), C.byref(sur.mlp), float(dt),
                      int(subcycles), self.spec.brillouin_iters, int(has_pabs), self._st.surrogate_tol,
                      sur.fallbacks.vp)

    def _consume_window(self):
        """_consume_multiphysics_window! (Solver.jl:373-383)."""
        window = self.n - self.last_mp_n
        if window <= 0:
            return
        self._film_update(self.dt * window, self.spec.subcycles, 1)
        self.last_mp_n = self.n

    def flush_multiphysics(self):
        """flush_multiphysics! (Solver.jl:387-390)."""
        if self.multiphysics_enabled and self.film is not None:
            self._consume_window()
        return self

    def freeze_multiphysics(self):
        self.multiphysics_enabled = False
        return self

    def enable_multiphysics(self):
        self.multiphysics_enabled = True
        self.last_mp_n = self.n
        return self

    def relax_step(self, dt: float, subcycles: Optional[int] = None):
        """relax_step! (Solver.jl:548-555): 4TM+LLB with no heating.  Like the reference it
        leaves n and t alone; the relax phase driver (phases.run_relax) advances them."""
        if self.film is None:
            return self
        self._film_update(float(dt), int(subcycles or self.spec.subcycles), 0)
        return self

    # ---- host access -------------------------------------------------------------------------
    def synchronize(self):
        _lib.call("synchronize")

    def set_field(self, name: str, values: np.ndarray) -> None:
        """Overwrite one field volume from a host (nx, ny, nz) array (adapt_backend)."""
        a = np.asarray(values, dtype=self.spec.field_dtype)
        self.f[name.lower()].upload(a.reshape(-1, order="F") if a.ndim == 3 else a)

    def field(self, name: str) -> np.ndarray:
        """One field volume as a host (nx, ny, nz) array (to_host)."""
        nx, ny, nz = self.spec.shape
        return self.f[name.lower()].download(self.spec.field_dtype).reshape((nx, ny, nz), order="F")

    def film_state(self) -> dict:
        if self.film is None:
            return {}
        out = {k: self.film_bufs[k].download(self.spec.compute_dtype) for k in FILM_NAMES}
        for k in ("U_abs", "window_sum", "peak"):
            out[k] = self.film_bufs[k].download(np.float64)
        return out

    def magnetization_snapshot(self) -> dict:
        """magnetization_snapshot (Solver.jl:448-467)."""
        fs = self.film_state()
        out = {k: fs.get(k, np.zeros(0)).astype(np.float64) for k in FILM_NAMES[4:]}
        out["material_cells"] = self.spec.film_cells.copy()
        return out

    def apply_magnetization(self, data: dict):
        """apply_magnetization! (Solver.jl:474-483)."""
        if self.film is None:
            raise ValueError("state has no MagnetizationState")
        for key in FILM_NAMES[4:]:
            src = np.asarray(data[key], dtype=np.float64).astype(self.spec.compute_dtype)
            if len(src) != self.spec.n_film:
                raise ValueError(f"magnetization {key} length mismatch")
            self.film_bufs[key].upload(src)
        return self

    def set_film(self, key: str, values: np.ndarray):
        dt = np.float64 if key in ("U_abs", "window_sum", "peak") else self.spec.compute_dtype
        self.film_bufs[key].upload(np.asarray(values).astype(dt))

    def psi(self, which: str, s: int) -> np.ndarray:
        """Raw storage of CPML slab s (0..11, order xy_lo .. zy_hi) of the H ('h') or
        D ('d') update; BF16 / binary16 kinds come back as their uint16 bit patterns."""
        buf = (self.psi_h if which == "h" else self.psi_d)[s]
        dt = {"f32": np.float32, "f64": np.float64, "bf16": np.uint16, "f16": np.uint16}[self.spec.psi_dtype]
        return buf.download(dt) if buf.ptr else np.zeros(0, dt)

    def ade_state(self, comp: int) -> dict:
        """P, J and E_old of ADE list comp; reduced kinds come back as uint16 bit patterns
        of P*scale and J*(dt/2)*scale."""
        sd = self.spec.ade_state_dtype
        b = self.ade_bufs[comp]
        dt = {"f32": np.float32, "f64": np.float64, "bf16": np.uint16, "f16": np.uint16}[sd]
        return dict(P=b["P"].download(dt), J=b["J"].download(dt), E_old=b["E_old"].download(self.spec.compute_dtype))

    def configure_probe_mode(self, lambda0: float = 532e-9, reset_ade: bool = True,
                             freeze_magnetization: bool = True, enable_magneto_optic: Optional[bool] = None):
        """configure_probe_mode! (Solver.jl:416-443): refit the poles at the probe
        wavelength, zero the ADE / MO histories and freeze the magnetisation."""
        from .grid import FDTDParams
        if self.film is None:
            return self
        spec = self.spec
        gd = spec.gd
        p = FDTDParams()
        CT = spec.compute_dtype
        if enable_magneto_optic is None:
            enable_magneto_optic = self.mo is not None
        if self.ade is not None:
            dp = dispersion.poles_array(dispersion.build_probe_poles(self.dt, p.eps0, gd), CT)
            if len(dp) != len(spec.diag_poles):
                raise ValueError("probe pole set must have the pump set's pole count")
            self.poles_dev.upload(dp)
        if enable_magneto_optic:
            mp_ = dispersion.poles_array(dispersion.build_probe_mo_poles(self.dt, p.eps0, gd), CT)
            if self.mo is None:
                if spec.ade is None:
                    raise ValueError("MO gyration needs the per-component lists")
                from .spec import ADEList, position_map
                comp_b = {"x": 0, "y": 1, "z": 2}[spec.mo_pair[1]]
                base = spec.ade[comp_b]
                spec.mo = ADEList(base.idx.copy(), base.fill.copy(), base.inv.copy())
                spec.mo_pos = position_map(spec.mo.idx, spec.film_cells)
                self._build_mo(mp_)
            else:
                self.mo_poles_dev.upload(mp_)
        else:
            self.mo = None
        self._build_state()
        if reset_ade:
            if self.ade_bufs is not None:
                for b in self.ade_bufs:
                    for k in ("P", "J", "E_old"):
                        _lib.call("memset", b[k].vp, 0, C.c_size_t(b[k].nbytes))
            if self.mo is not None:
                for k in ("Pa", "Ja", "Pb", "Jb", "Ea_old", "Eb_old"):
                    b = self.mo_bufs[k]
                    _lib.call("memset", b.vp, 0, C.c_size_t(b.nbytes))
        if freeze_magnetization:
            self.freeze_multiphysics()
        return self

    def free(self):
        """free_state! (Solver.jl:521-546)."""
        for b in self._bufs:
            b.free()
        for d in (getattr(self, "film_bufs", {}), getattr(self, "mo_bufs", {})):
            for b in d.values():
                b.free()
        for b in (getattr(self, "lut_tm", None), getattr(self, "lut_re", None)):
            if b is not None:
                b.free()
        self._bufs = []
