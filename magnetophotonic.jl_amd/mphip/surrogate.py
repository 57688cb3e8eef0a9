"""Learned stand-in for the sub-stepped 4TM + LLB film integrator (host side).

The network (MLP 11 -> 64 -> 64 -> 6, SiLU, BF16 MFMA with FP32 accumulation) and its
training run on the GPU through libmphip:

  mp_surrogate_dataset   training rows (features, rates of the plain integrator) from a
                         film state, without changing it
  mp_mlp_backward        forward + backward over a batch (deterministic gradient)
  mp_mlp_adam            Adam on the FP32 master parameters
  mp_mlp_pack            BF16 MFMA fragments from the FP32 parameters
  mp_surrogate_step      the film update: network rates, checked against one plain
                         substep, plain fallback where the check fails

It stands in for the substep loop of multiphysics_kernel! (src/fdtd/Kernels.jl:362-412);
the reference has no learned component.  Features, targets and the parameter layout
are defined in include/mphip.h (mp_mlp).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional

import numpy as np

from . import _lib, abi
from .state import DeviceBuffer, FDTDState

N_IN, N_OUT = 11, 6
NPARAM, NNORM, PACKED = abi.MLP_NPARAM, abi.MLP_NNORM, abi.MLP_PACKED_BYTES
OW1, OB1, OW2, OB2, OW3, OB3 = 0, 1024, 1088, 5184, 5248, 6272


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights on the real rows/columns, zero biases and padding."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)

    def glorot(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, (fo, fi))

    W1 = np.zeros((64, 16))
    W1[:, :N_IN] = glorot(64, N_IN)
    p[OW1:OB1] = W1.reshape(-1)
    p[OW2:OB2] = glorot(64, 64).reshape(-1)
    W3 = np.zeros((16, 64))
    W3[:N_OUT] = glorot(N_OUT, 64)
    p[OW3:OB3] = W3.reshape(-1)
    return p


def make_norm(X: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Normalisation constants (x_mean, x_inv_std, y_mean, y_std, y_inv_std) from rows;
    a feature that does not vary gets inv_std 0."""
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    xm, xs = X.mean(0), X.std(0)
    ym, ys = T.mean(0), T.std(0)
    norm = np.zeros(NNORM, np.float32)
    norm[0:N_IN] = xm
    const = xs <= 1e-12 * np.maximum(np.abs(xm), 1.0)
    norm[16:16 + N_IN] = np.where(const, 0.0, 1.0 / np.where(const, 1.0, xs))
    ys = np.where(ys > 0, ys, 1.0)
    norm[32:32 + N_OUT] = ym
    norm[48:48 + N_OUT] = ys
    norm[64:64 + N_OUT] = 1.0 / ys
    return norm


class Surrogate:
    """Device-resident network: FP32 master parameters, BF16 fragments, normalisation,
    Adam moments and the fallback counter."""

    def __init__(self, params: Optional[np.ndarray] = None, norm: Optional[np.ndarray] = None, seed: int = 0):
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters")
        self.params = DeviceBuffer.from_array(p)
        n = np.zeros(NNORM, np.float32) if norm is None else np.asarray(norm, np.float32)
        if norm is None:
            n[16:16 + N_IN] = 1.0
            n[48:48 + N_OUT] = 1.0
            n[64:64 + N_OUT] = 1.0
        self.norm = DeviceBuffer.from_array(n)
        self.packed = DeviceBuffer.zeros(PACKED)
        self.grad = DeviceBuffer.zeros(NPARAM * 4)
        self.m1 = DeviceBuffer.zeros(NPARAM * 4)
        self.m2 = DeviceBuffer.zeros(NPARAM * 4)
        self.loss_dev = DeviceBuffer.zeros(8)
        self.fallbacks = DeviceBuffer.zeros(8)
        self.mlp = abi.Mlp(self.params.ptr, self.packed.ptr, self.norm.ptr)
        self.adam_steps = 0
        self.pack()

    # ---- parameters ------------------------------------------------------------------------
    def pack(self) -> None:
        _lib.call("mlp_pack", C.byref(self.mlp))

    def set_norm(self, norm: np.ndarray) -> None:
        self.norm.upload(np.asarray(norm, np.float32))

    def get_params(self) -> np.ndarray:
        return self.params.download(np.float32)

    def get_norm(self) -> np.ndarray:
        return self.norm.download(np.float32)

    def set_params(self, p: np.ndarray) -> None:
        self.params.upload(np.asarray(p, np.float32))
        self.pack()

    # ---- compute ---------------------------------------------------------------------------
    def forward(self, X: DeviceBuffer, B: int) -> DeviceBuffer:
        """Raw rates (B x 6) for raw features X (B x 11), both device FP32."""
        Y = DeviceBuffer(max(B, 1) * N_OUT * 4)
        _lib.call("mlp_forward", C.byref(self.mlp), X.vp, Y.vp, int(B))
        return Y

    def backward(self, X_ptr: int, T_ptr: int, B: int) -> None:
        """dloss/dparams into self.grad and the loss into self.loss_dev."""
        _lib.call("mlp_backward", C.byref(self.mlp), C.c_void_p(X_ptr), C.c_void_p(T_ptr), int(B),
                  self.grad.vp, self.loss_dev.vp)

    def loss(self) -> float:
        return float(self.loss_dev.download(np.float64)[0])

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, wd: float = 0.0):
        self.adam_steps += 1
        _lib.call("mlp_adam", self.params.vp, self.grad.vp, self.m1.vp, self.m2.vp, self.adam_steps,
                  float(lr), float(beta1), float(beta2), float(eps), float(wd))
        self.pack()

    def train(self, data: "Dataset", steps: int, batch: int = 65536, lr: float = 3e-3, lr_end: float = 1e-4,
              log_every: int = 0) -> list:
        """Adam over contiguous batches of the (shuffled) dataset, cosine learning-rate
        decay from lr to lr_end; returns the loss at every log_every-th step."""
        if data.rows < batch:
            batch = data.rows
        nb = data.rows // batch
        hist = []
        for s in range(int(steps)):
            k = s % nb
            self.backward(data.X.ptr + k * batch * N_IN * 4, data.T.ptr + k * batch * N_OUT * 4, batch)
            a = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * s / max(steps - 1, 1)))
            self.adam(a)
            if log_every and (s % log_every == 0 or s == steps - 1):
                hist.append((s, self.loss()))
        return hist

    def evaluate(self, data: "Dataset", rows: Optional[int] = None, batch: int = 65536) -> float:
        """Mean normalised squared error over the first `rows` rows (gradient discarded)."""
        rows = data.rows if rows is None else min(rows, data.rows)
        tot, cnt = 0.0, 0
        for r0 in range(0, rows, batch):
            b = min(batch, rows - r0)
            self.backward(data.X.ptr + r0 * N_IN * 4, data.T.ptr + r0 * N_OUT * 4, b)
            tot += self.loss() * b
            cnt += b
        return tot / max(cnt, 1)

    def reset_fallbacks(self) -> None:
        self.fallbacks.upload(np.zeros(1, np.int64))

    def fallback_count(self) -> int:
        return int(self.fallbacks.download(np.int64)[0])

    def free(self) -> None:
        for b in (self.params, self.norm, self.packed, self.grad, self.m1, self.m2, self.loss_dev, self.fallbacks):
            b.free()


class Dataset:
    """Training rows in device memory: X (rows x 11) raw features, T (rows x 6) raw rates."""

    def __init__(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        if X.shape[0] != T.shape[0]:
            raise ValueError("row count mismatch")
        self.rows = X.shape[0]
        self.X = DeviceBuffer.from_array(X)
        self.T = DeviceBuffer.from_array(T)

    def free(self) -> None:
        self.X.free()
        self.T.free()


def film_rows(st: FDTDState, dt: float, subcycles: int, has_pabs: bool, stride: int = 1):
    """Training rows from the film of `st` (not modified): host (X, T)."""
    n = st.spec.n_film
    rows = (n + stride - 1) // stride
    Xd = DeviceBuffer(rows * N_IN * 4)
    Td = DeviceBuffer(rows * N_OUT * 4)
    _lib.call("surrogate_dataset", C.byref(st.film), st.gd_arr.ctypes.data_as(C.c_void_p), C.byref(st.lut),
              float(dt), int(subcycles), st.spec.brillouin_iters, 1 if has_pabs else 0, int(stride), Xd.vp, Td.vp,
              rows)
    X = Xd.download(np.float32).reshape(rows, N_IN)
    T = Td.download(np.float32).reshape(rows, N_OUT)
    Xd.free()
    Td.free()
    return X, T


def collect_trajectory(st: FDTDState, pump_steps: int, every: int, relax_steps: int = 0, relax_dt: float = 5e-15,
                       relax_subcycles: int = 10, relax_every: int = 10, stride: int = 1):
    """Rows along a plain run: every `every` pump steps the film state just before the
    film update of that step (labelled with the step's own dt, subcycles and absorbed
    power), then every `relax_every` relax steps (no heating, relax_dt)."""
    Xs, Ts = [], []
    sub = st.spec.subcycles
    done = 0
    while done < pump_steps:
        k = min(every - 1, pump_steps - done - 1)
        if k > 0:
            st.run(k)
            done += k
        # one step with the film update held back, so the window holds this step's power
        st.freeze_multiphysics()
        st.step()
        done += 1
        X, T = film_rows(st, st.dt * (st.n - st.last_mp_n), sub, True, stride)
        Xs.append(X)
        Ts.append(T)
        st.multiphysics_enabled = True
        st.flush_multiphysics()
    st.flush_multiphysics()
    for r in range(relax_steps):
        if r % relax_every == 0:
            X, T = film_rows(st, relax_dt, relax_subcycles, False, stride)
            Xs.append(X)
            Ts.append(T)
        st.relax_step(relax_dt, subcycles=relax_subcycles)
    return np.concatenate(Xs), np.concatenate(Ts)


def shuffled(X: np.ndarray, T: np.ndarray, seed: int = 0):
    perm = np.random.default_rng(seed).permutation(X.shape[0])
    return X[perm], T[perm]


# ---- the pump + relax accuracy study ---------------------------------------------------------
# A 128 x 128 x 64 copy of the coupled pump geometry (film 10 cells of 2 nm, source 60 nm above
# it, CPML 8): 1500 pump steps at the given fluence, then relax steps of 5 fs with 10 plain
# substeps.  At 60 mJ/cm^2 about 5 % of the film switches, so the film-mean magnetisation after
# the relax phase is sensitive to small errors in the film update.
ACCURACY_CASE = dict(nx=128, ny=128, nz=64, npml=8, src_z=60e-9, n_planes=0)
FILM_KEYS = ("Te", "Tl", "Ts_TM", "Ts_RE", "m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z")


def relax_rows(fluences, pump_steps: int = 1500, relax_steps: int = 600, relax_dt: float = 5e-15,
               relax_subcycles: int = 10, every: int = 5, stride: int = 4):
    """Training rows of the relax phase (no heating, relax_dt, plain integrator) after the
    accuracy case's pump at each fluence: every `every`-th relax step, every `stride`-th cell."""
    from . import configs
    Xs, Ts = [], []
    for fl in fluences:
        spec, _ = configs.coupled_pump(dtype="f32", steps=pump_steps, fluence_mj_cm2=float(fl), **ACCURACY_CASE)
        st = FDTDState(spec)
        st.run(pump_steps)
        st.flush_multiphysics()
        for r in range(relax_steps):
            if r % every == 0:
                X, T = film_rows(st, relax_dt, relax_subcycles, False, stride)
                Xs.append(X)
                Ts.append(T)
            st.relax_step(relax_dt, subcycles=relax_subcycles)
        st.free()
    return np.concatenate(Xs), np.concatenate(Ts)


def train_relax_surrogate(X, T, steps: int = 30000, seed: int = 1) -> "Surrogate":
    X, T = shuffled(X, T, 0)
    sur = Surrogate(norm=make_norm(X[:200000], T[:200000]), seed=seed)
    data = Dataset(X, T)
    sur.train(data, steps, batch=65536, lr=3e-3, lr_end=1e-5)
    data.free()
    return sur


def pump_relax_film(dtype: str = "f32", fluence: float = 60.0, pump_steps: int = 1500,
                    relax_steps: int = 600, relax_dt: float = 5e-15, relax_subcycles: int = 10) -> dict:
    """Film state of the accuracy case after pump + relax, with the plain integrator on the
    GPU (dtype 'f64' gives the double-precision reference of the same path)."""
    from . import configs
    spec, _ = configs.coupled_pump(dtype=dtype, steps=pump_steps, fluence_mj_cm2=fluence, **ACCURACY_CASE)
    st = FDTDState(spec)
    st.run(pump_steps)
    st.flush_multiphysics()
    for _ in range(relax_steps):
        st.relax_step(relax_dt, subcycles=relax_subcycles)
    film = st.film_state()
    st.free()
    return film


def relax_comparison(sur: "Surrogate", ref_film: dict, fluence: float = 60.0, pump_steps: int = 1500,
                     relax_steps: int = 600, relax_dt: float = 5e-15, relax_subcycles: int = 10,
                     tols=(1e-3, 1e-4), plain_subcycles=(10, 1, 2, 3)) -> dict:
    """Relax phase of the accuracy case from one FP32 pump state, by the plain integrator
    with several substep counts and by the surrogate at several tolerances: film-mean
    |dm| against ref_film (max over both sublattices), wall time and fallback fraction."""
    import time

    from . import _lib, configs
    spec, _ = configs.coupled_pump(dtype="f32", steps=pump_steps, fluence_mj_cm2=fluence, **ACCURACY_CASE)
    st = FDTDState(spec)
    st.run(pump_steps)
    st.flush_multiphysics()
    start = st.film_state()
    n = spec.n_film

    def one(sub, tol=None):
        for k in FILM_KEYS:
            st.set_film(k, start[k])
        if tol is not None:
            st.attach_surrogate(sur, tol)
            sur.reset_fallbacks()
        _lib.call("synchronize")
        t0 = time.perf_counter()
        for _ in range(relax_steps):
            st.relax_step(relax_dt, subcycles=sub)
        _lib.call("synchronize")
        ms = 1e3 * (time.perf_counter() - t0)
        f = st.film_state()
        fb = sur.fallback_count() / (n * relax_steps) if tol is not None else 0.0
        st.detach_surrogate()
        err = max(abs(float(f[k].astype(np.float64).mean() - np.asarray(ref_film[k], np.float64).mean()))
                  for k in ("m_TM_x", "m_RE_x"))
        return dict(film_mean_dm=err, relax_ms=round(ms, 2), fallback=round(fb, 4),
                    switched=float(np.mean(f["m_TM_x"] < 0)))

    out = {f"plain_{s}sub": one(s) for s in plain_subcycles}
    for tol in tols:
        out[f"surrogate_tol{tol:g}"] = one(relax_subcycles, tol)
    out["switched_ref"] = float(np.mean(np.asarray(ref_film["m_TM_x"]) < 0))
    st.free()
    return out


# ---- the storage + surrogate forms on the accuracy case ----------------------------------------
def pump_relax_rows(fluences, pump_steps: int = 1500, pump_every: int = 25, relax_steps: int = 600,
                    relax_every: int = 5, stride: int = 4, relax_dt: float = 5e-15, relax_subcycles: int = 10):
    """Training rows of both phases of the accuracy case at each fluence (plain FP32 runs):
    every `pump_every`-th pump step and every `relax_every`-th relax step, every `stride`-th
    film cell."""
    from . import configs
    Xs, Ts = [], []
    for fl in fluences:
        spec, _ = configs.coupled_pump(dtype="f32", steps=pump_steps, fluence_mj_cm2=float(fl), **ACCURACY_CASE)
        st = FDTDState(spec)
        X, T = collect_trajectory(st, pump_steps, pump_every, relax_steps, relax_dt, relax_subcycles, relax_every,
                                  stride)
        st.free()
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def film_mean_dm(film: dict, ref_film: dict) -> float:
    """Final film-mean |dm| against a reference film state, the larger of the two sublattices."""
    return max(abs(float(np.asarray(film[k], np.float64).mean() - np.asarray(ref_film[k], np.float64).mean()))
               for k in ("m_TM_x", "m_RE_x"))


def accuracy_run(kind: Optional[str] = None, sur: Optional["Surrogate"] = None, tol: float = 1e-4,
                 fluence: float = 60.0, pump_steps: int = 1500, relax_steps: int = 600, relax_dt: float = 5e-15,
                 relax_subcycles: int = 10) -> dict:
    """The accuracy case (pump + relax) in one form: CPML psi and ADE state stored as `kind`
    ('bf16' / 'f16'; None keeps FP32) and the film update by `sur` in both phases (None: the
    plain integrator).  Returns the final film state, the GPU wall time of the whole case and
    the fraction of film-cell updates that fell back to the plain integrator."""
    import time

    from . import _lib, configs
    kw = dict(psi_precision=kind, ade_state_precision=kind) if kind else {}
    spec, _ = configs.coupled_pump(dtype="f32", steps=pump_steps, fluence_mj_cm2=fluence, **ACCURACY_CASE, **kw)
    st = FDTDState(spec)
    if sur is not None:
        st.attach_surrogate(sur, tol)
        sur.reset_fallbacks()
    _lib.call("synchronize")
    t0 = time.perf_counter()
    st.run(pump_steps)
    st.flush_multiphysics()
    for _ in range(relax_steps):
        st.relax_step(relax_dt, subcycles=relax_subcycles)
    _lib.call("synchronize")
    ms = 1e3 * (time.perf_counter() - t0)
    updates = spec.n_film * (pump_steps // max(spec.multiphysics_every, 1) + relax_steps)
    fb = sur.fallback_count() / updates if sur is not None else 0.0
    film = st.film_state()
    st.free()
    return dict(film=film, ms=ms, fallback=fb)


def accuracy_forms(kinds=("bf16", "f16"), tol: float = 1e-4, fluences=(35.0, 50.0, 70.0, 85.0)) -> dict:
    """Film-mean |dm| of the accuracy case at 60 mJ/cm^2 against the same path in FP64 on
    the GPU, for plain FP32, for the surrogate with FP32 storage and for the surrogate with
    each reduced storage kind.  One network serves both phases; it is trained on the GPU on
    rows of both phases of runs at the training fluences (60 is held out)."""
    import time
    t0 = time.perf_counter()
    X, T = pump_relax_rows(fluences)
    sur = train_relax_surrogate(X, T)
    rows = int(X.shape[0])
    del X, T
    ref = pump_relax_film("f64")
    out = {}

    def entry(r):
        return dict(film_mean_dm=film_mean_dm(r["film"], ref), case_ms=round(r["ms"], 1),
                    fallback=round(r["fallback"], 4), switched=float(np.mean(np.asarray(r["film"]["m_TM_x"]) < 0)))

    out["plain"] = entry(accuracy_run(None))
    out["surrogate"] = entry(accuracy_run(None, sur, tol))
    for k in kinds:
        out[f"surrogate+{k}"] = entry(accuracy_run(k, sur, tol))
    out["switched_ref"] = float(np.mean(np.asarray(ref["m_TM_x"]) < 0))
    out["train_rows"] = rows
    out["tol"] = tol
    out["study_s"] = round(time.perf_counter() - t0, 1)
    sur.free()
    return out


# ---- the C2 pump trajectories ------------------------------------------------------------------
C2_SHAPE = (256, 256, 192)
C2_FLUENCES = (2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 14.0, 16.0)


def c2_rows(fluences=C2_FLUENCES, steps: int = 4000, every: int = 100, stride: int = 8):
    """Training rows from plain runs of the 256 x 256 x 192 coupled pump at each fluence
    (mJ/cm^2): `steps` steps with the pulse peak in the middle, every `every`-th step the film
    state before that step's film update, every `stride`-th film cell."""
    from . import configs
    nx, ny, nz = C2_SHAPE
    Xs, Ts = [], []
    for fl in fluences:
        spec, _ = configs.coupled_pump(nx=nx, ny=ny, nz=nz, steps=steps, fluence_mj_cm2=float(fl), n_planes=0)
        st = FDTDState(spec)
        X, T = collect_trajectory(st, steps, every, stride=stride)
        st.free()
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)
