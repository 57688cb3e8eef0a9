"""Probe read-out, polarimetry and transmission monitors: device reductions + host algebra.

`ProbeReadout` mirrors src/sim/ProbeReadout.jl (struct :11-75, constructor :79-107,
trace sampling :111-112, _probe_init! :123-170, device record :274-301, probe_shot
:362-478).  The running DFT of the pre and post x-planes is mp_probe_dft_accumulate
(recorded inside the step driver while `run` drives the state), the mode-weighted
traces are mp_probe_trace_plane into device slots, and probe_shot's plane reductions
(_plane_poynting, _plane_net_flux, _jones_overlap) are one mp_probe_shot_reduce call:
only 4*nb + 10 doubles cross to the host per shot.  The scalar algebra after the
reductions (T, R, A, per-bin T/R with the relative floor, Jones angles, trace moments)
is `shot_from_reductions`, which needs no device.

`Polarimetry`, `Transmission` and `Reflection` mirror the 3-D record! and monitor_data of
src/sim/Monitors.jl (:386-403, :461-485, :608-635) with the plane sums on the device
(mp_plane_overlap, mp_flux_plane) and their series kept in device slots.

Movie frames (_probe_frame_due / _record_probe_frame!, ProbeReadout.jl:183-215) are staged
on the device: at each due step one mp_probe_frames launch writes |Ey| and |Ez| (float) of
the z slice and the pre and post x-planes into frame-last device cubes (the layout of
_stack_frames, :172-181); full chunks of frames are downloaded, and probe_shot returns the
six cubes with the frame indices and times (:465-475).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from .grid import FDTDParams

_EPS64 = float(np.finfo(np.float64).eps)
_COMPONENTS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}
_AXES = {"x": 0, "y": 1, "z": 2}


# ---- host algebra (no device) ----------------------------------------------------------

def probe_jones_angles_deg(ey: complex, ez: complex) -> dict:
    """Rotation from the Ez axis and ellipticity of the Jones pair (Ey, Ez), in degrees
    (probe_jones_angles_deg, src/physics/Polarimetry.jl:5-14)."""
    ey, ez = complex(ey), complex(ez)
    den = abs(ey) ** 2 + abs(ez) ** 2
    if not math.isfinite(den) or den <= 0.0:
        return dict(rotation_deg=math.nan, ellipticity_deg=math.nan)
    cross = ey * ez.conjugate()
    theta = 0.5 * math.atan2(2.0 * cross.real, abs(ez) ** 2 - abs(ey) ** 2)
    s3 = min(max(2.0 * cross.imag / den, -1.0), 1.0)
    chi = 0.5 * math.asin(s3)
    return dict(rotation_deg=theta * 180.0 / math.pi, ellipticity_deg=chi * 180.0 / math.pi)


def energy_balance(*, incident: float, reflected: float, transmitted: float) -> dict:
    """T, R and the closure A = 1 - R - T (energy_balance, Polarimetry.jl:17-23)."""
    inc = float(incident)
    R = float(reflected) / inc
    T = float(transmitted) / inc
    return dict(T=T, R=R, A=1.0 - R - T)


def trace_moments(t: Sequence[float], y: Sequence[float]) -> dict:
    """|y|^2-weighted mean time and rms width, in fs (_trace_moments, ProbeReadout.jl:312-321)."""
    t = np.asarray(t, np.float64)
    if t.size == 0:
        return dict(delay_fs=math.nan, rms_fs=math.nan)
    w = np.abs(np.asarray(y, np.float64)) ** 2
    sw = float(np.sum(w))
    if sw <= 0:
        return dict(delay_fs=math.nan, rms_fs=math.nan)
    mu = float(np.sum(t * w)) / sw
    sig = math.sqrt(max(float(np.sum((t - mu) ** 2 * w)) / sw, 0.0))
    return dict(delay_fs=mu * 1e15, rms_fs=sig * 1e15)


def unpack_reductions(red: np.ndarray, nb: int) -> dict:
    """Split the 4*nb + 10 values of mp_probe_shot_reduce."""
    red = np.asarray(red, np.float64)
    if red.shape != (4 * nb + 10,):
        raise ValueError(f"expected {4 * nb + 10} reduced values, got {red.shape}")
    sx_pre = red[0:2 * nb:2] + 1j * red[1:2 * nb:2]
    sx_post = red[2 * nb:4 * nb:2] + 1j * red[2 * nb + 1:4 * nb:2]
    j = red[4 * nb + 2:4 * nb + 10]
    return dict(Sx_pre=sx_pre, Sx_post=sx_post, net_pre=float(red[4 * nb]), net_post=float(red[4 * nb + 1]),
                far_Ey=complex(j[0], j[1]), far_Ez=complex(j[2], j[3]),
                kerr_Ey=complex(j[4], j[5]), kerr_Ez=complex(j[6], j[7]))


def shot_from_reductions(red: Optional[dict], nb: int, mode_norm: float, *, incident_energy=None,
                         incident_sx=None, absorbed_energy=None) -> dict:
    """The scalar part of probe_shot (ProbeReadout.jl:374-420) from the plane reductions
    (`unpack_reductions` output; None = an uninitialised read-out, all sums zero)."""
    if red is None:
        red = dict(Sx_pre=np.zeros(nb, complex), Sx_post=np.zeros(nb, complex), net_pre=0.0, net_post=0.0,
                   far_Ey=0j, far_Ez=0j, kerr_Ey=0j, kerr_Ez=0j)
        mode_norm = 1.0
    sx_pre, sx_post = np.asarray(red["Sx_pre"], complex), np.asarray(red["Sx_post"], complex)
    net_pre, net_post = float(red["net_pre"]), float(red["net_post"])
    # bare-waveguide reference normalisation: a reference shot's incident is its own
    # pre-plane net; later shots pass the reference's values in
    incident = net_pre if incident_energy is None else float(incident_energy)
    sx_inc = sx_pre.copy() if incident_sx is None else np.asarray(incident_sx, complex).copy()
    denom = incident if abs(incident) > 0.0 else _EPS64
    reflected = incident - net_pre
    transmitted = net_post
    T = transmitted / denom
    R = reflected / denom
    absorbed = max(incident - transmitted - reflected, 0.0) if absorbed_energy is None else float(absorbed_energy)
    A = absorbed / denom
    inc_re = sx_inc.real
    peak = float(np.max(np.abs(inc_re))) if inc_re.size else 0.0
    den_floor = 1.0e-12 * peak
    T_omega = np.full(nb, math.nan)
    R_omega = np.full(nb, math.nan)
    for q in range(nb):
        den = inc_re[q]
        if abs(den) > den_floor:
            T_omega[q] = sx_post[q].real / den
            R_omega[q] = (den - sx_pre[q].real) / den
    inv = 1.0 / mode_norm
    jf = (red["far_Ey"] * inv, red["far_Ez"] * inv)
    jk = (red["kerr_Ey"] * inv, red["kerr_Ez"] * inv)
    far = probe_jones_angles_deg(*jf)
    ker = probe_jones_angles_deg(*jk)
    return dict(
        T=T, R=R, A=A, T_plus_R_plus_A=T + R + A, T_omega=T_omega, R_omega=R_omega,
        theta_faraday_deg=far["rotation_deg"], theta_kerr_deg=ker["rotation_deg"],
        ellipticity_faraday_deg=far["ellipticity_deg"], ellipticity_kerr_deg=ker["ellipticity_deg"],
        jones_faraday_Ey_re_im=[jf[0].real, jf[0].imag], jones_faraday_Ez_re_im=[jf[1].real, jf[1].imag],
        jones_kerr_Ey_re_im=[jk[0].real, jk[0].imag], jones_kerr_Ez_re_im=[jk[1].real, jk[1].imag],
        Sx_incident_re=sx_inc.real.copy(), Sx_incident_im=sx_inc.imag.copy(),
        Sx_pre_net_re=sx_pre.real.copy(), Sx_pre_net_im=sx_pre.imag.copy(),
        Sx_post_net_re=sx_post.real.copy(), Sx_post_net_im=sx_post.imag.copy(), Sx_inc=sx_inc,
        incident_energy_J=incident, transmitted_energy_J=transmitted, reflected_energy_J=reflected,
        absorbed_energy_J=absorbed, net_energy_pre_J=net_pre, net_energy_post_J=net_post,
        throughput_loss=1.0 - T,
    )


def mo_consistency_metrics(plus_M, zero_M, minus_M, tolerance: float = 5.0e-3) -> dict:
    """+M / 0 / -M consistency of T+R+A and the Faraday angle
    (mo_consistency_metrics, ProbeReadout.jl:480-506).  Shots are dicts."""
    def tra(s):
        return float(s["T_plus_R_plus_A"]) if "T_plus_R_plus_A" in s else float(s["T"] + s["R"] + s["A"])
    plus, zero, minus = tra(plus_M), tra(zero_M), tra(minus_M)
    if not all(math.isfinite(v) for v in (plus, zero, minus)):
        raise ValueError(f"MO consistency totals must be finite: +M={plus}, 0={zero}, -M={minus}")
    spread = max(plus, zero, minus) - min(plus, zero, minus)
    theta = [float(plus_M["theta_faraday_deg"]), float(zero_M["theta_faraday_deg"]),
             float(minus_M["theta_faraday_deg"])]
    if not all(math.isfinite(v) for v in theta):
        raise ValueError("MO consistency Faraday angles must be finite")
    pedestal = 0.5 * (theta[0] + theta[2])
    tol = float(tolerance)
    return dict(TRA_plus_M=plus, TRA_zero_M=zero, TRA_minus_M=minus, even_in_M_violation_plus_minus=plus - minus,
                TRA_spread_plus_zero_minus=spread, theta_faraday_deg_plus_zero_minus=theta,
                pedestal_avg_vs_zeroM=[pedestal, theta[1]], tolerance=tol, conservative=spread <= tol)


def probe_contrast(initial: dict, switched: dict) -> dict:
    """INITIAL - SWITCHED deltas of two shots (probe_contrast, ProbeReadout.jl:508-525)."""
    d = {}
    for k in ("T", "R", "A", "theta_faraday_deg", "theta_kerr_deg", "ellipticity_faraday_deg",
              "ellipticity_kerr_deg", "group_delay_fs", "pulse_broadening_fs"):
        d["delta_" + k] = initial[k] - switched[k]
    d["initial_T_R_A"] = [initial["T"], initial["R"], initial["A"]]
    d["switched_T_R_A"] = [switched["T"], switched["R"], switched["A"]]
    d["initial_T_plus_R_plus_A"] = initial["T_plus_R_plus_A"]
    d["switched_T_plus_R_plus_A"] = switched["T_plus_R_plus_A"]
    return d


def default_bins(lambda0: float, spectrum_bins: int = 5):
    """omega_probe * (1 + linspace(-0.04, 0.04, n)) and the bin nearest omega_probe
    (ProbeReadout.jl:89-94; argmin takes the first of equal distances)."""
    omega_probe = 2.0 * math.pi * FDTDParams().c0 / float(lambda0)
    n = int(spectrum_bins)
    if n < 2:
        # a range from -0.04 to 0.04 needs two points; the reference raises as well
        raise ValueError("spectrum_bins must be at least 2 (pass omega_bins for a single bin)")
    return omega_probe * (1.0 + np.linspace(-0.04, 0.04, n)), omega_probe


def center_bin(bins, omega_probe: float) -> int:
    return int(np.argmin(np.abs(np.asarray(bins, np.float64) - omega_probe)))


def polarimetry_summary(t, ey, ez, omega: float) -> dict:
    """Single-frequency phasors sum(e_i * exp(i*omega*t_i)) and their Jones angles
    (monitor_data(::Polarimetry), Monitors.jl:626-635)."""
    t = np.asarray(t, np.float64)
    if t.size == 0:
        return dict(t=t, ey=np.asarray(ey), ez=np.asarray(ez), rotation_deg=math.nan, ellipticity_deg=math.nan)
    ph = np.exp(1j * float(omega) * t)
    Ey = complex(np.sum(np.asarray(ey, np.float64) * ph))
    Ez = complex(np.sum(np.asarray(ez, np.float64) * ph))
    ang = probe_jones_angles_deg(Ey, Ez)
    return dict(t=t, ey=np.asarray(ey), ez=np.asarray(ez), rotation_deg=ang["rotation_deg"],
                ellipticity_deg=ang["ellipticity_deg"])


def tr_summary(t, values, modal: bool, incident: float, sign: float = 1.0) -> dict:
    """Energy and ratio of a Transmission / Reflection series (_tr_summary, Monitors.jl:612-621)."""
    t = np.asarray(t, np.float64)
    v = np.asarray(values, np.float64)
    dt = float(t[1] - t[0]) if t.size > 1 else 0.0
    energy = float(np.sum(v ** 2) * dt) if modal else float(np.sum(v) * dt)
    ratio = sign * energy / max(float(incident), _EPS64)
    return dict(t=t, values=v, modal=modal, energy=energy, ratio=ratio)


# ---- device-backed monitors ---------------------------------------------------------------

class _Slots:
    """A growing device series of `width` doubles per record; records stay on the device
    until `values()` downloads them."""

    def __init__(self, width: int, chunk: int = 1024):
        self.width, self.chunk = int(width), int(chunk)
        self.host = []
        self.buf = None
        self.count = 0

    def next_ptr(self):
        from .state import DeviceBuffer
        if self.buf is None:
            self.buf = DeviceBuffer.zeros(8 * self.width * self.chunk)
            self.count = 0
        if self.count == self.chunk:
            self.host.append(self.buf.download(np.float64).reshape(self.chunk, self.width))
            self.count = 0
        p = C.c_void_p(self.buf.ptr + 8 * self.width * self.count)
        self.count += 1
        return p

    def values(self) -> np.ndarray:
        parts = list(self.host)
        if self.buf is not None and self.count:
            parts.append(self.buf.download(np.float64, self.width * self.count).reshape(self.count, self.width))
        return np.concatenate(parts, 0) if parts else np.zeros((0, self.width))

    def free(self):
        if self.buf is not None:
            self.buf.free()
            self.buf = None


class _FrameStore:
    """Probe movie frames on the device: six frame-last float32 cubes, (nx, ny) slices of
    Ey, Ez at z = kc and (ny, nz) slices at the pre and post x-planes, in chunks of `chunk`
    frames; a full chunk is downloaded before the next frame is staged."""

    KEYS = ("Ey_xy", "Ez_xy", "Ey_yz_pre", "Ez_yz_pre", "Ey_yz_post", "Ez_yz_post")

    def __init__(self, nx: int, ny: int, nz: int, chunk: int = 32):
        self.dims = {k: ((nx, ny) if k.endswith("_xy") else (ny, nz)) for k in self.KEYS}
        self.chunk = int(chunk)
        self.count = 0
        self.host = {k: [] for k in self.KEYS}
        self.bufs = None

    def _download(self, k, frames):
        d1, d2 = self.dims[k]
        a = self.bufs[k].download(np.float32, d1 * d2 * frames)
        return a.reshape((d1, d2, frames), order="F")

    def next_ptrs(self):
        """Device pointers of the next frame's six planes (mp_probe_frames order)."""
        from .state import DeviceBuffer
        if self.bufs is None:
            self.bufs = {k: DeviceBuffer.zeros(4 * d1 * d2 * self.chunk) for k, (d1, d2) in self.dims.items()}
            self.count = 0
        if self.count == self.chunk:
            for k in self.KEYS:
                self.host[k].append(self._download(k, self.chunk))
            self.count = 0
        ptrs = []
        for k in self.KEYS:
            d1, d2 = self.dims[k]
            ptrs.append(C.c_void_p(self.bufs[k].ptr + 4 * d1 * d2 * self.count))
        self.count += 1
        return ptrs

    def frames(self) -> dict:
        out = {}
        for k in self.KEYS:
            parts = list(self.host[k])
            if self.bufs is not None and self.count:
                parts.append(self._download(k, self.count))
            out[k] = np.concatenate(parts, axis=2) if parts else np.zeros((0, 0, 0), np.float32)
        return out

    def free(self):
        if self.bufs is not None:
            for b in self.bufs.values():
                b.free()
            self.bufs = None


def _plane_index(state, axis: int, position) -> int:
    from .sources import source_index
    g = state.spec.grid
    return source_index((g.x, g.y, g.z)[axis], position)


class ProbeReadout:
    """Gold-standard probe read-out on the device (ProbeReadout.jl:11-478)."""

    def __init__(self, pre, post, mode=None, lambda0: float = 532e-9, omega_bins=None, spectrum_bins: int = 5,
                 trace_stride: int = 10, frame_target: int = 0, frame_skip: int = 0, store_frames: bool = True,
                 probe_amplitude: float = 1.0, probe_tau: float = 24e-15, frame_indices=(),
                 slice_z_position: float = math.nan, final_step: int = 0):
        self.x_pre, self.x_post = pre, post
        # the frame plan and its stride (ProbeReadout.jl:98-102)
        plan = [int(v) for v in frame_indices]
        skip = int(frame_skip)
        if skip <= 0 and len(plan) > 1:
            skip = max(1, plan[1] - plan[0])
        self.frame_target, self.frame_skip, self.store_frames = int(frame_target), skip, bool(store_frames)
        self.frame_plan, self.frame_ptr = plan, 0
        self.slice_z_position = float(slice_z_position)
        self.frame_indices, self.frame_times = [], []
        self.mode = mode
        self.lambda0 = float(lambda0)
        if omega_bins is None:
            bins, omega_probe = default_bins(lambda0, spectrum_bins)
        else:
            bins = np.asarray(omega_bins, np.float64).reshape(-1)
            omega_probe = 2.0 * math.pi * FDTDParams().c0 / float(lambda0)
        self.omega_bins = np.asarray(bins, np.float64)
        self.center_bin = center_bin(self.omega_bins, omega_probe)
        self.trace_stride = int(trace_stride)
        self.probe_amplitude, self.probe_tau = float(probe_amplitude), float(probe_tau)
        self.final_step = int(final_step)
        self.initialized = False
        self.t_trace = []

    # -- set-up -----------------------------------------------------------------------
    def _init(self, state):
        from .monitors import DFTPlanes, plane_area
        from .state import DeviceBuffer
        g = state.spec.grid
        self.i_pre = _plane_index(state, 0, self.x_pre)
        self.i_post = _plane_index(state, 0, self.x_post)
        ny, nz = g.y.n, g.z.n
        self.area = plane_area(g, 0)
        have_mode = self.mode is not None and np.shape(self.mode) == (ny, nz)
        self.mode_w = (np.asarray(self.mode, np.float64).reshape(-1, order="F") if have_mode
                       else np.ones(ny * nz))
        norm = float(np.sum(np.abs(self.mode_w) * self.area))
        self.mode_norm = norm if norm > 0.0 else 1.0
        self.planes = DFTPlanes(state, [(0, self.i_pre), (0, self.i_post)], self.omega_bins)
        self.area_dev = DeviceBuffer.from_array(self.area)
        self.mode_w_dev = DeviceBuffer.from_array(self.mode_w)
        self.red_dev = DeviceBuffer.zeros(8 * (4 * len(self.omega_bins) + 10))
        self.traces = _Slots(4)
        # the xy movie slice: the domain's middle z plane unless a position is given (:197-198)
        self.kc = (-(-g.z.n // 2) - 1) if math.isnan(self.slice_z_position) else _plane_index(
            state, 2, self.slice_z_position)
        self.frames_store = _FrameStore(g.x.n, ny, nz)
        self.state = state
        self.initialized = True

    def _ensure(self, state):
        if not self.initialized:
            self._init(state)
        elif state is not self.state:
            raise ValueError("a ProbeReadout records one state")

    # -- recording ----------------------------------------------------------------------
    def due(self, n: int) -> bool:
        return True

    def trace_due(self, n: int) -> bool:
        """n == 1, every trace_stride-th step, and the final step (ProbeReadout.jl:111-112)."""
        return n == 1 or n % self.trace_stride == 0 or (self.final_step > 0 and n == self.final_step)

    def frame_due(self, n: int) -> bool:
        """_probe_frame_due (ProbeReadout.jl:183-191): the next planned step, else every
        frame_skip-th step (or every ceil(n / frame_target)-th) until frame_target frames."""
        if self.frame_target <= 0:
            return False
        if self.frame_plan:
            return self.frame_ptr < len(self.frame_plan) and n == self.frame_plan[self.frame_ptr]
        skip = self.frame_skip if self.frame_skip > 0 else max(1, -(-max(n, 1) // self.frame_target))
        return n % skip == 0 and len(self.frame_indices) < self.frame_target

    def _record_frame(self, state):
        """_record_probe_frame! (ProbeReadout.jl:193-215) with the planes staged on the device."""
        if not self.frame_due(state.n):
            return
        if self.frame_plan:
            self.frame_ptr += 1
        self.frame_indices.append(int(state.n))
        self.frame_times.append(float(state.t))
        if not self.store_frames:
            return
        from . import _lib
        _lib.call("probe_frames", C.byref(state.yee), self.kc, self.i_pre, self.i_post,
                  *self.frames_store.next_ptrs())

    def _record_trace(self, state):
        from . import _lib
        _lib.call("probe_trace_plane", C.byref(state.yee), 0, self.i_pre, self.i_post, self.mode_w_dev.vp,
                  self.area_dev.vp, self.traces.next_ptr())
        self.t_trace.append(float(state.t))

    def record(self, state) -> "ProbeReadout":
        """record! after one step taken outside `run` (ProbeReadout.jl:303-307)."""
        self._ensure(state)
        self.planes.record(state)
        if self.trace_due(state.n):
            self._record_trace(state)
        self._record_frame(state)
        return self

    def run(self, state, steps: int) -> "ProbeReadout":
        """`steps` steps of `state` with the read-out recorded after each: the DFT inside the
        step driver, the traces and movie frames between driver calls at the due steps."""
        self._ensure(state)
        state.attach_dft(self.planes)
        try:
            end = state.n + int(steps)
            while state.n < end:
                n = state.n + 1
                while n < end and not (self.trace_due(n) or self.frame_due(n)):
                    n += 1
                state.run(n - state.n)
                if self.trace_due(state.n):
                    self._record_trace(state)
                self._record_frame(state)
        finally:
            state.detach_dft()
        return self

    # -- read-out ----------------------------------------------------------------------
    def reductions(self) -> Optional[dict]:
        """The device plane reductions of the current accumulators (None before the first record)."""
        if not self.initialized:
            return None
        from . import _lib
        nb = len(self.omega_bins)
        arr = self.planes.arr
        g = self.state.spec.grid
        _lib.call("probe_shot_reduce", C.byref(arr[0]), C.byref(arr[1]), g.y.n, g.z.n, nb, self.center_bin,
                  self.mode_w_dev.vp, self.area_dev.vp, self.red_dev.vp)
        return unpack_reductions(self.red_dev.download(np.float64), nb)

    def trace_values(self) -> dict:
        v = self.traces.values() / self.mode_norm if self.initialized else np.zeros((0, 4))
        return dict(t=np.asarray(self.t_trace, np.float64), Ey_pre=v[:, 0], Ez_pre=v[:, 1], Ey_post=v[:, 2],
                    Ez_post=v[:, 3])

    def probe_shot(self, state_label: str = "probe", is_reference: bool = False, incident_energy=None,
                   incident_sx=None, normalization_method: str = "gold_standard_bare_waveguide_reference_netflux",
                   dt: float = math.nan, steps_probe: int = 0, absorbed_energy=None) -> dict:
        """probe_shot (ProbeReadout.jl:362-478)."""
        nb = len(self.omega_bins)
        red = self.reductions()
        out = shot_from_reductions(red, nb, self.mode_norm if self.initialized else 1.0,
                                   incident_energy=incident_energy, incident_sx=incident_sx,
                                   absorbed_energy=absorbed_energy)
        tr = self.trace_values()
        inp = trace_moments(tr["t"], tr["Ez_pre"])
        trans = trace_moments(tr["t"], tr["Ez_post"])
        g = self.state.spec.grid if self.initialized else None
        out.update(
            group_delay_fs=trans["delay_fs"] - inp["delay_fs"],
            pulse_broadening_fs=trans["rms_fs"] - inp["rms_fs"],
            input_pulse_rms_fs=inp["rms_fs"], transmitted_pulse_rms_fs=trans["rms_fs"],
            Ey_trace_pre=tr["Ey_pre"], Ez_trace_pre=tr["Ez_pre"], Ey_trace_post=tr["Ey_post"],
            Ez_trace_post=tr["Ez_post"], trace_time_s=tr["t"], omega_bins_rad_s=self.omega_bins.copy(),
            x_pre_um=(float(g.x.centers[self.i_pre]) * 1e6 if g is not None else math.nan),
            x_post_um=(float(g.x.centers[self.i_post]) * 1e6 if g is not None else math.nan),
            is_reference=1 if is_reference else 0, state_label=str(state_label),
            normalization_method=str(normalization_method), dt=float(dt), lambda_nm=self.lambda0 * 1e9,
            probe_amplitude_V_m=self.probe_amplitude, probe_tau_fs=self.probe_tau * 1e15,
            probe_trace_stride=self.trace_stride, steps_probe=int(steps_probe),
            probe_frame_target=self.frame_target, probe_frame_skip=self.frame_skip,
            probe_frame_count=len(self.frame_indices), probe_frame_indices=list(self.frame_indices),
            probe_frame_times_s=list(self.frame_times),
            frames=(self.frames_store.frames() if self.initialized
                    else {k: np.zeros((0, 0, 0), np.float32) for k in _FrameStore.KEYS}),
        )
        return out

    def free(self):
        if self.initialized:
            self.planes.free()
            for b in (self.area_dev, self.mode_w_dev, self.red_dev):
                b.free()
            self.traces.free()
            self.frames_store.free()
            self.initialized = False


class _PlaneSeries:
    def __init__(self, width: int):
        self.t = []
        self.slots = _Slots(width)

    def due(self, n: int) -> bool:
        return True

    def free(self):
        self.slots.free()


class Polarimetry(_PlaneSeries):
    """Plane-mean (no mode) or mode-overlap Ey, Ez per record (record!(::Polarimetry),
    Monitors.jl:476-485 with _mean_plane_components :461-473 and _mode_overlap :350-368)."""

    def __init__(self, at=None, axis: str = "x", mode=None, omega: float = 2 * math.pi * 299792458.0 / 532e-9):
        super().__init__(2)
        self.at, self.axis, self.mode, self.omega = at, axis, mode, float(omega)
        self._ready = False

    def _init(self, state):
        from .monitors import plane_area
        from .state import DeviceBuffer
        g = state.spec.grid
        if self.mode is None:
            # the mean is always over an x-plane, the last one when no position is given
            self.ax = 0
            self.idx = g.x.n - 1 if self.at is None else _plane_index(state, 0, self.at)
            self.w_dev = None
            self.ncell = g.y.n * g.z.n
        else:
            self.ax = _AXES[self.axis]
            self.idx = _plane_index(state, self.ax, self.at)
            self.w_dev = DeviceBuffer.from_array(_mode_weight(g, self.ax, self.mode) * plane_area(g, self.ax))
            self.ncell = None
        self._ready = True

    def record(self, state) -> "Polarimetry":
        from . import _lib
        if not self._ready:
            self._init(state)
        p = self.slots.next_ptr()
        w = self.w_dev.vp if self.w_dev is not None else None
        _lib.call("plane_overlap", C.byref(state.yee), 1, self.ax, self.idx, w, p)
        _lib.call("plane_overlap", C.byref(state.yee), 2, self.ax, self.idx, w, C.c_void_p(p.value + 8))
        self.t.append(float(state.t))
        return self

    def data(self) -> dict:
        v = self.slots.values()
        if self.ncell is not None:
            v = v / max(self.ncell, 1)
        return polarimetry_summary(self.t, v[:, 0], v[:, 1], self.omega)

    def free(self):
        super().free()
        if getattr(self, "w_dev", None) is not None:
            self.w_dev.free()


class Transmission(_PlaneSeries):
    """Signed plane flux (no mode) or the mode overlap of one component per record
    (record!(::Transmission) / record!(::Reflection), Monitors.jl:386-403, 3-D branch)."""

    _sign = 1.0

    def __init__(self, at, axis: str = "x", component: str = "Ez", mode=None, incident: float = 1.0):
        super().__init__(1)
        self.at, self.axis, self.component, self.mode = at, axis, component, mode
        self.incident = float(incident)
        self._ready = False

    def _init(self, state):
        from .monitors import plane_area
        from .state import DeviceBuffer
        g = state.spec.grid
        self.ax = _AXES[self.axis]
        self.idx = _plane_index(state, self.ax, self.at)
        self.area_dev = DeviceBuffer.from_array(plane_area(g, self.ax))
        self.w_dev = (DeviceBuffer.from_array(_mode_weight(g, self.ax, self.mode) * plane_area(g, self.ax))
                      if self.mode is not None else None)
        self._ready = True

    def record(self, state):
        from . import _lib
        if not self._ready:
            self._init(state)
        p = self.slots.next_ptr()
        if self.mode is None:
            _lib.call("flux_plane", C.byref(state.yee), self.ax, self.idx, self.area_dev.vp, p, 0)
        else:
            _lib.call("plane_overlap", C.byref(state.yee), _COMPONENTS[self.component], self.ax, self.idx,
                      self.w_dev.vp, p)
        self.t.append(float(state.t))
        return self

    def values(self) -> np.ndarray:
        v = self.slots.values()[:, 0]
        return -v if (self.mode is None and self._sign < 0) else v

    def data(self) -> dict:
        s = tr_summary(self.t, self.values(), self.mode is not None, self.incident)
        s["transmission" if self._sign > 0 else "reflection"] = s["ratio"]
        return s

    def free(self):
        super().free()
        if getattr(self, "area_dev", None) is not None:
            self.area_dev.free()
        if getattr(self, "w_dev", None) is not None:
            self.w_dev.free()


class Reflection(Transmission):
    """As Transmission with the plane flux negated (Monitors.jl:396-403)."""

    _sign = -1.0


def _mode_weight(grid, axis: int, mode) -> np.ndarray:
    """The reference's mode array on a plane — (ny, nz) for x, (nx, nz) for y, (nx, ny) for
    z — in this library's plane order (x: (ny, nz), y: (nz, nx), z: (nx, ny)), flattened."""
    m = np.asarray(mode, np.float64)
    n = ((grid.y.n, grid.z.n), (grid.x.n, grid.z.n), (grid.x.n, grid.y.n))[axis]
    if m.shape != n:
        raise ValueError(f"mode shape {m.shape} does not match the plane {n}")
    if axis == 1:
        m = m.T
    return m.reshape(-1, order="F")
