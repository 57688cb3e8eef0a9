"""ORACLE step driver -- test infrastructure, not product code.

Runs the reference's step sequence (src/fdtd/Solver.jl:283-383 `_step_typed!`,
`_consume_multiphysics_window!`; Solver.jl:548-555 `relax_step!`) over host numpy
arrays with the C restatement in libmp_oracle.so.  Only tests, the smoke check and the
benchmark's CPU-baseline leg import this module.

It consumes the same host-side StateSpec as the GPU state, so both start from identical
inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmp_oracle.so")

_lib = None


def build() -> str:
    """Compile the oracle (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib(abi):
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        abi.bind(h, "orc_", abi.SIGNATURES)
        h.orc_llb_step.restype = None
        h.orc_llb_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_int32,
                                   C.c_void_p]
        h.orc_update_4tm_cell.restype = None
        h.orc_update_4tm_cell.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                          C.c_double]
        h.orc_brillouin_fn.restype = C.c_double
        h.orc_brillouin_fn.argtypes = [C.c_double, C.c_double]
        h.orc_lut_lookup.restype = C.c_double
        h.orc_lut_lookup.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32]
        vp, P = C.c_void_p, C.POINTER
        h.orc_surrogate_dataset.restype = C.c_int
        h.orc_surrogate_dataset.argtypes = [P(abi.Film), vp, P(abi.Lut), C.c_double, C.c_int32, C.c_int32,
                                            C.c_int32, C.c_int64, vp, vp, C.c_int64]
        h.orc_surrogate_apply.restype = C.c_int
        h.orc_surrogate_apply.argtypes = [P(abi.Film), vp, P(abi.Lut), vp, C.c_double, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_double, P(C.c_int64)]
        for name in ("orc_store_reduced", "orc_load_reduced"):
            getattr(h, name).restype = None
            getattr(h, name).argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_float]
        _lib = h
    return _lib


# storage kind codes of include/mphip.h
_KIND = {"f32": 0, "f64": 1, "bf16": 2, "f16": 3}


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class OracleState:
    """Host twin of mphip.FDTDState driven by the C oracle."""

    FIELD_NAMES = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")
    FILM_NAMES = ("Te", "Tl", "Ts_TM", "Ts_RE", "m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z")

    def __init__(self, spec):
        from mphip import abi  # struct layouts only
        from mphip.cpml import PSI_AXIS
        from mphip.sources import ModeSource

        self.abi = abi
        self.L = lib(abi)
        self.spec = spec
        self.n = 0
        self.t = 0.0
        self.last_mp_n = 0
        self.dt = spec.dt
        self.multiphysics_enabled = spec.multiphysics_enabled
        nx, ny, nz = spec.shape
        FT, CT = spec.field_dtype, spec.compute_dtype
        code = {np.dtype(np.float32): abi.MP_F32, np.dtype(np.float64): abi.MP_F64}
        self.fcode, self.ccode = code[FT], code[CT]
        self.f = {k: np.zeros(nx * ny * nz, FT) for k in self.FIELD_NAMES}
        self.inv_eps = [a.copy() for a in spec.inv_eps]
        self.idc = [a.copy() for a in spec.inv_d_cell]
        self.idd = [a.copy() for a in spec.inv_d_dual]
        y = abi.Yee()
        y.nx, y.ny, y.nz = nx, ny, nz
        y.field_dtype, y.compute_dtype = self.fcode, self.ccode
        for k in self.FIELD_NAMES:
            setattr(y, k, self.f[k].ctypes.data)
        y.inv_eps_x, y.inv_eps_y, y.inv_eps_z = (a.ctypes.data for a in self.inv_eps)
        y.inv_d_cell_x, y.inv_d_cell_y, y.inv_d_cell_z = (a.ctypes.data for a in self.idc)
        y.inv_d_dual_x, y.inv_d_dual_y, y.inv_d_dual_z = (a.ctypes.data for a in self.idd)
        self.yee = y
        cs = spec.cpml
        cp = abi.Cpml()
        cp.npml_x, cp.npml_y, cp.npml_z = cs.widths
        psi_np = {"f32": np.float32, "f64": np.float64, "bf16": np.uint16, "f16": np.uint16}[spec.psi_dtype]
        cp.psi_dtype = _KIND[spec.psi_dtype]
        cp.psi_scale = spec.psi_scale
        self.coef = []
        for d, ax in enumerate(cs.axes):
            arrs = (ax.inv_kappa.copy(), ax.a.copy(), ax.b.copy())
            self.coef.append(arrs)
            cp.inv_kappa[d], cp.a[d], cp.b[d] = (a.ctypes.data for a in arrs)
        self.psi_h = [np.zeros(cs.slab_cells(PSI_AXIS[s]), psi_np) for s in range(12)]
        self.psi_d = [np.zeros(cs.slab_cells(PSI_AXIS[s]), psi_np) for s in range(12)]
        for s in range(12):
            cp.psi_h[s] = self.psi_h[s].ctypes.data if self.psi_h[s].size else None
            cp.psi_d[s] = self.psi_d[s].ctypes.data if self.psi_d[s].size else None
        self.cpml = cp

        src = spec.source
        if isinstance(src, ModeSource):
            import math
            self.mode_profile = np.asarray(src.profile, np.float64).reshape(-1, order="F").astype(CT)
            from mphip.sources import source_index
            self.mode_i = source_index(spec.grid.x, src.position)
            self.mode_hcorr = src.neff / math.sqrt(spec.params.mu0 / spec.params.eps0)

        self.ade = None
        if spec.ade is not None:
            np_ = len(spec.diag_poles) // 3
            self.poles = spec.diag_poles.copy()
            self.ade, self.ade_arr = [], []
            for lst in spec.ade:
                n = len(lst.idx)
                sd = spec.ade_state_dtype
                SD = np.uint16 if sd in ("bf16", "f16") else CT
                arr = dict(P=np.zeros(n * np_, SD), J=np.zeros(n * np_, SD), E_old=np.zeros(n, CT),
                           idx=lst.idx.copy(), fill=lst.fill.copy(), inv=lst.inv.copy())
                a = abi.Ade()
                a.n, a.npoles, a.dtype = n, np_, self.ccode
                a.state_dtype, a.state_scale = _KIND[sd], spec.ade_state_scale
                a.P, a.J, a.E_old = _ptr(arr["P"]), _ptr(arr["J"]), _ptr(arr["E_old"])
                a.idx, a.fill, a.inv_eps = _ptr(arr["idx"]), _ptr(arr["fill"]), _ptr(arr["inv"])
                a.poles = _ptr(self.poles)
                self.ade.append(a)
                self.ade_arr.append(arr)

        self.film = None
        if spec.has_film:
            n = spec.n_film
            self.fa = {k: spec.film_init[k].copy() for k in self.FILM_NAMES}
            self.fa["cells"] = spec.film_cells.copy()
            for c, pos in zip("xyz", spec.pos):
                self.fa["pos_" + c] = pos.copy()
            for k in ("U_abs", "window_sum", "peak"):
                self.fa[k] = np.zeros(n)
            fm = abi.Film()
            fm.n, fm.dtype = n, self.ccode
            for k, a in self.fa.items():
                setattr(fm, k, a.ctypes.data if a.size else None)
            self.film = fm
            self.gd_arr = np.ascontiguousarray(spec.gd.as_array(np.float64))
            tm, re_, T_min, inv_dT, nT = spec.lut
            self.lut_tm, self.lut_re = tm.copy(), re_.copy()
            lut = abi.Lut()
            lut.tm, lut.re, lut.T_min, lut.inv_dT, lut.n, lut.dtype = (
                self.lut_tm.ctypes.data, self.lut_re.ctypes.data, T_min, inv_dT, nT, self.ccode)
            self.lut = lut

        self.mo = None
        if spec.mo is not None:
            n = len(spec.mo.idx)
            np_ = len(spec.mo_poles) // 3
            self.mo_poles = spec.mo_poles.copy()
            self.ma = {k: np.zeros(n * np_, CT) for k in ("Pa", "Ja", "Pb", "Jb")}
            self.ma.update(Ea_old=np.zeros(n, CT), Eb_old=np.zeros(n, CT), idx=spec.mo.idx.copy(),
                           fill=spec.mo.fill.copy(), mpos=spec.mo_pos.copy())
            m = abi.Mo()
            m.n, m.npoles, m.dtype = n, np_, self.ccode
            m.Pa_from_b, m.Ja_from_b = _ptr(self.ma["Pa"]), _ptr(self.ma["Ja"])
            m.Pb_from_a, m.Jb_from_a = _ptr(self.ma["Pb"]), _ptr(self.ma["Jb"])
            m.E_old_a, m.E_old_b = _ptr(self.ma["Ea_old"]), _ptr(self.ma["Eb_old"])
            m.idx, m.fill, m.mpos, m.poles = (_ptr(self.ma["idx"]), _ptr(self.ma["fill"]),
                                              _ptr(self.ma["mpos"]), _ptr(self.mo_poles))
            self.mo = m
            a, b = spec.mo_pair
            ci = {"x": 0, "y": 1, "z": 2}
            self.mo_e = (self.f["e" + a], self.f["e" + b])
            self.mo_ie = (self.inv_eps[ci[a]], self.inv_eps[ci[b]])

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what}: {self.L.orc_last_error().decode()}")

    def _inject(self):
        from mphip.sources import (AXIS_CODE, COMPONENT_CODE, ModeSource, PlaneSource, SoftSource,
                                   axis_weights)
        src = self.spec.source
        y = C.byref(self.yee)
        if src is None:
            return
        if isinstance(src, SoftSource):
            i, j, k = src.index
            self._check(self.L.orc_inject_soft_3d(y, COMPONENT_CODE[src.component], i, j, k,
                                                  src.pulse.value(self.t)), "inject_soft")
        elif isinstance(src, ModeSource):
            self._check(self.L.orc_inject_mode_x_3d(y, COMPONENT_CODE[src.component], self.mode_i,
                                                    _ptr(self.mode_profile), src.pulse.value(self.t),
                                                    self.mode_hcorr), "inject_mode")
        elif isinstance(src, PlaneSource):
            g = self.spec.grid
            axis = {"x": g.x, "y": g.y, "z": g.z}[src.axis]
            w = axis_weights(axis, src.position)
            i0, w0 = w[0]
            i1, w1 = w[1] if len(w) > 1 else (-1, 0.0)
            self._check(self.L.orc_inject_plane_3d(y, COMPONENT_CODE[src.component], AXIS_CODE[src.axis],
                                                   i0, w0, i1, w1, src.pulse.value(self.t)), "inject_plane")

    def step(self):
        spec = self.spec
        y, cp = C.byref(self.yee), C.byref(self.cpml)
        self._check(self.L.orc_update_H_3d(y, cp, self.dt, spec.params.mu0), "H")
        self._inject()
        self._check(self.L.orc_update_E_3d(y, cp, self.dt), "E")
        if self.ade is not None:
            for comp, a in zip(("ex", "ey", "ez"), self.ade):
                self._check(self.L.orc_patch_E_dispersive(_ptr(self.f[comp]), C.byref(a), self.fcode, self.dt), "ADE")
        if self.mo is not None and self.film is not None:
            self._check(self.L.orc_patch_E_mo_gyration(
                _ptr(self.mo_e[0]), _ptr(self.mo_e[1]), C.byref(self.mo), _ptr(self.mo_ie[0]),
                _ptr(self.mo_ie[1]), self.fcode, _ptr(self.fa["m_TM_x"]), _ptr(self.fa["m_RE_x"]),
                float(spec.gd["Q_voigt_TM"]), float(spec.gd["Q_voigt_RE"]), self.dt), "MO")
        self.n += 1
        self.t += self.dt
        if self.film is not None:
            hw = 1 if spec.has_ade_work() else 0
            ax = C.byref(self.ade[0]) if hw else None
            ay = C.byref(self.ade[1]) if hw else None
            az = C.byref(self.ade[2]) if hw else None
            self._check(self.L.orc_pabs_accumulate(C.byref(self.film), y, ax, ay, az, hw, spec.pabs_factor(),
                                                   self.dt), "pabs")
            if self.multiphysics_enabled and spec.multiphysics_every > 0 and self.n % spec.multiphysics_every == 0:
                self._consume()
        return self

    def _consume(self):
        window = self.n - self.last_mp_n
        if window <= 0:
            return
        self._check(self.L.orc_multiphysics_step(C.byref(self.film), _ptr(self.gd_arr), C.byref(self.lut),
                                                 self.dt * window, self.spec.subcycles,
                                                 self.spec.brillouin_iters, 1), "multiphysics")
        self.last_mp_n = self.n

    def flush_multiphysics(self):
        if self.multiphysics_enabled and self.film is not None:
            self._consume()
        return self

    def relax_step(self, dt, subcycles=None):
        if self.film is None:
            return self
        self._check(self.L.orc_multiphysics_step(C.byref(self.film), _ptr(self.gd_arr), C.byref(self.lut),
                                                 float(dt), int(subcycles or self.spec.subcycles),
                                                 self.spec.brillouin_iters, 0), "relax")
        return self

    def run(self, steps):
        for _ in range(int(steps)):
            self.step()
        return self

    def set_field(self, name, values):
        a = np.asarray(values, dtype=self.f[name.lower()].dtype)
        self.f[name.lower()][...] = a.reshape(-1, order="F") if a.ndim == 3 else a

    def field(self, name):
        nx, ny, nz = self.spec.shape
        return self.f[name.lower()].reshape((nx, ny, nz), order="F").copy()

    def film_state(self):
        if self.film is None:
            return {}
        return {k: self.fa[k].copy() for k in self.FILM_NAMES + ("U_abs", "window_sum", "peak")}

    def set_film(self, key, values):
        self.fa[key][...] = np.asarray(values).astype(self.fa[key].dtype)


class OracleDFT:
    """Host twin of mphip.monitors.DFTPlanes over the oracle (orc_probe_dft_accumulate)."""

    def __init__(self, ref, planes, omega_bins):
        from mphip import abi
        from mphip.monitors import _plane_dims
        self.ref, self.planes, self.nb = ref, [(int(a), int(i)) for a, i in planes], len(omega_bins)
        self.omega = np.asarray(omega_bins, np.float64)
        self.phase = np.zeros(2 * self.nb)
        self.arrs = []
        self.desc = (abi.Plane * max(1, len(self.planes)))()
        for p, (axis, idx) in enumerate(self.planes):
            n1, n2 = _plane_dims(ref.spec.shape, axis)
            a = {k: np.zeros(2 * n1 * n2 * self.nb) for k in ("ea", "eb", "ha", "hb")}
            a["energy"] = np.zeros(n1 * n2)
            self.arrs.append((a, n1, n2))
            d = self.desc[p]
            d.axis, d.index = axis, idx
            d.dft_ea, d.dft_eb, d.dft_ha, d.dft_hb = (a[k].ctypes.data for k in ("ea", "eb", "ha", "hb"))
            d.energy = a["energy"].ctypes.data

    def record(self, ref=None, t=None, dt=None):
        ref = ref or self.ref
        rc = ref.L.orc_probe_dft_accumulate(C.byref(ref.yee), self.desc, len(self.planes),
                                            self.omega.ctypes.data_as(C.c_void_p),
                                            self.phase.ctypes.data_as(C.c_void_p), self.nb,
                                            ref.t if t is None else float(t), ref.dt if dt is None else float(dt))
        assert rc == 0

    def download(self, p):
        a, n1, n2 = self.arrs[p]
        out = {k: (a[k][0::2] + 1j * a[k][1::2]).reshape((n1, n2, self.nb), order="F")
               for k in ("ea", "eb", "ha", "hb")}
        out["energy"] = a["energy"].reshape((n1, n2), order="F")
        return out

    def raw(self, p):
        """The flat accumulator arrays of plane p (interleaved re/im), as the device holds them."""
        return self.arrs[p][0]


class OracleProbe:
    """Host twin of mphip.probe.ProbeReadout: DFT and traces recorded by the oracle after
    every oracle step, plane reductions by orc_probe_shot_reduce; the scalar algebra is the
    library's shot_from_reductions applied to these reductions."""

    def __init__(self, ref, readout):
        from mphip.monitors import plane_area
        from mphip.sources import source_index
        g = ref.spec.grid
        self.ref, self.ro = ref, readout
        self.i_pre = source_index(g.x, readout.x_pre)
        self.i_post = source_index(g.x, readout.x_post)
        self.area = plane_area(g, 0)
        ny, nz = g.y.n, g.z.n
        m = readout.mode
        have = m is not None and np.shape(m) == (ny, nz)
        self.mode_w = np.asarray(m, np.float64).reshape(-1, order="F") if have else np.ones(ny * nz)
        norm = float(np.sum(np.abs(self.mode_w) * self.area))
        self.mode_norm = norm if norm > 0.0 else 1.0
        self.dft = OracleDFT(ref, [(0, self.i_pre), (0, self.i_post)], readout.omega_bins)
        self.t_trace, self.traces = [], []

    def record(self):
        ref = self.ref
        self.dft.record(ref)
        if self.ro.trace_due(ref.n):
            out = np.zeros(4)
            rc = ref.L.orc_probe_trace_plane(C.byref(ref.yee), 0, self.i_pre, self.i_post, _ptr(self.mode_w),
                                             _ptr(self.area), _ptr(out))
            assert rc == 0
            self.t_trace.append(ref.t)
            self.traces.append(out / self.mode_norm)

    def run(self, steps):
        for _ in range(int(steps)):
            self.ref.step()
            self.record()

    def reductions_raw(self):
        nb = self.dft.nb
        out = np.zeros(4 * nb + 10)
        g = self.ref.spec.grid
        rc = self.ref.L.orc_probe_shot_reduce(C.byref(self.dft.desc[0]), C.byref(self.dft.desc[1]), g.y.n, g.z.n,
                                              nb, self.ro.center_bin, _ptr(self.mode_w), _ptr(self.area), _ptr(out))
        assert rc == 0
        return out


def plane_overlap(ref, component, axis, index, weight=None):
    """orc_plane_overlap on the oracle state's fields."""
    out = np.zeros(1)
    w = None if weight is None else np.ascontiguousarray(weight, np.float64)
    rc = ref.L.orc_plane_overlap(C.byref(ref.yee), int(component), int(axis), int(index), _ptr(w), _ptr(out))
    assert rc == 0
    return float(out[0])


def film_average(ref):
    """orc_film_average on the oracle state's film: the 13 sums."""
    out = np.zeros(13)
    assert ref.L.orc_film_average(C.byref(ref.film), _ptr(out)) == 0
    return out


def switching_sums(ref, core_fraction=0.5):
    """orc_switching_metrics on the oracle state's film: the 13 counts and sums."""
    out = np.zeros(13)
    assert ref.L.orc_switching_metrics(C.byref(ref.film), float(core_fraction), _ptr(out)) == 0
    return out
