"""The Julia extension's call sequences, replayed through ctypes against the oracle.

There is no Julia toolchain in this image, so the two ways ext/MagnetoPhotonicHIPExt.jl
drives libmphip are replayed here with the struct contents it builds:

- the step driver: `step!` -> `_step_typed!(::HIPBackend)` -> one `mp_run_steps` call per
  step with an mp_state whose Yee struct carries no material index (the three inv_eps
  volumes), no DFT record and the fused film pass where the lists coincide; `run!` -> one
  call for all the steps;
- the launchers: the reference's `_step_typed!` sequence (Solver.jl:283-369), each launch
  the ext's `_ka_*` method with its struct contents: the H update's Yee struct without
  inv_eps or dual metrics, the E update's without cell metrics, the soft source's with the
  component's inv_eps in all three slots, the absorbed-power film struct with only the
  cell list, position maps and FP64 accumulators and ADE structs with only P, J, E_old and
  the list, the plane source through mp_inject_plane_3d.

Each runs 100 steps of a coupled case (and 100 of a plane-source case) against the oracle:
fields bit for bit where the film does not feed back into E, the coupled bar (1e-5
relative L2, m within 1e-4) where the gyration does.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d else float(np.linalg.norm(a - b))


class ExtLaunchers:
    """The ext's launcher path over a state's device buffers (the reference step order)."""

    def __init__(self, st):
        from mphip import abi
        self.st, self.abi, self.spec = st, abi, st.spec

    def _yee(self, inv_eps=(None, None, None), cell=False, dual=False):
        st, abi = self.st, self.abi
        y = abi.Yee()
        y.nx, y.ny, y.nz = st.spec.shape
        y.field_dtype, y.compute_dtype = st._fcode, st._ccode
        for k in FIELDS:
            setattr(y, k, st.f[k].ptr)
        y.inv_eps_x, y.inv_eps_y, y.inv_eps_z = (b.ptr if b is not None else None for b in inv_eps)
        if cell:
            y.inv_d_cell_x, y.inv_d_cell_y, y.inv_d_cell_z = (b.ptr for b in st.idc)
        if dual:
            y.inv_d_dual_x, y.inv_d_dual_y, y.inv_d_dual_z = (b.ptr for b in st.idd)
        return y

    def step(self):
        from mphip import _lib
        from mphip.sources import AXIS_CODE, COMPONENT_CODE, PlaneSource, SoftSource, axis_weights
        st, abi, spec = self.st, self.abi, self.spec
        dt = st.dt
        _lib.call("update_H_3d", C.byref(self._yee(cell=True)), C.byref(st.cpml), dt, spec.params.mu0)
        src = spec.source
        v = src.pulse.value(st.t)
        if isinstance(src, SoftSource):
            ie = st.inv_eps["xyz".index(src.component[1].lower())]
            i, j, k = src.index
            _lib.call("inject_soft_3d", C.byref(self._yee((ie, ie, ie))), COMPONENT_CODE[src.component], i, j, k, v)
        elif isinstance(src, PlaneSource):
            axis = {"x": spec.grid.x, "y": spec.grid.y, "z": spec.grid.z}[src.axis]
            w = axis_weights(axis, src.position)
            i1, w1 = w[1] if len(w) > 1 else (-1, 0.0)
            y = self._yee(tuple(st.inv_eps))
            y.compute_dtype = st._fcode          # the ext passes the field type here
            _lib.call("inject_plane_3d", C.byref(y), COMPONENT_CODE[src.component],
                      AXIS_CODE[src.axis], w[0][0], w[0][1], i1, w1, v)
        _lib.call("update_E_3d", C.byref(self._yee(tuple(st.inv_eps), dual=True)), C.byref(st.cpml), dt)
        if st.ade is not None:
            for c, comp in enumerate(("ex", "ey", "ez")):
                _lib.call("patch_E_dispersive", st.f[comp].vp, C.byref(st.ade[c]), st._fcode, dt)
        if st.mo is not None and st.film is not None:
            _lib.call("patch_E_mo_gyration", st._mo_ea.vp, st._mo_eb.vp, C.byref(st.mo), st._mo_iea.vp,
                      st._mo_ieb.vp, st._fcode, C.c_void_p(st.film.m_TM_x), C.c_void_p(st.film.m_RE_x),
                      st.Q_TM, st.Q_RE, dt)
        st.n += 1
        st.t += dt
        if st.film is not None:
            fb = st.film_bufs
            f = abi.Film()
            f.n, f.dtype = st.film.n, st.film.dtype
            f.cells, f.pos_x, f.pos_y, f.pos_z = fb["cells"].ptr, fb["pos_x"].ptr, fb["pos_y"].ptr, fb["pos_z"].ptr
            f.U_abs, f.window_sum, f.peak = fb["U_abs"].ptr, fb["window_sum"].ptr, fb["peak"].ptr
            hw = spec.has_ade_work() and st.ade is not None
            lists = []
            if hw:
                for c in range(3):
                    a, b = abi.Ade(), st.ade[c]
                    a.n, a.npoles, a.dtype, a.state_dtype, a.state_scale = b.n, b.npoles, b.dtype, b.dtype, 1.0
                    a.P, a.J, a.E_old, a.idx = b.P, b.J, b.E_old, b.idx
                    lists.append(a)
            ptrs = [C.byref(a) for a in lists] if hw else [None, None, None]
            _lib.call("pabs_accumulate", C.byref(f), C.byref(self._yee()), *ptrs, 1 if hw else 0,
                      spec.pabs_factor(), dt)
            if spec.multiphysics_enabled and spec.multiphysics_every > 0 and st.n % spec.multiphysics_every == 0:
                window = st.n - st.last_mp_n
                if window > 0:
                    _lib.call("multiphysics_step", C.byref(st.film), st.gd_arr.ctypes.data_as(C.c_void_p),
                              C.byref(st.lut), dt * window, spec.subcycles, spec.brillouin_iters, 1)
                    st.last_mp_n = st.n


def _cases():
    """The oracle's film maxima after 100 steps: Te 335.9 K (pump), 300.004 K (soft source
    three cells below the slab), 568.4 K (plane source)."""
    from mphip import configs, sources
    pump, _ = configs.coupled_pump(nx=40, ny=36, nz=64, steps=200, fluence_mj_cm2=12.0, npml=6, src_z=40e-9)
    small, _ = configs.small_coupled(dtype="f32", steps=100, amplitude=4e8, n=24, source_kind="soft")
    small.source = sources.SoftSource(small.source.pulse, "Ex", (12, 11, 5))
    plane, _ = configs.small_coupled(dtype="f32", steps=100, amplitude=4e8, n=24)
    zc = plane.grid.z.centers
    plane.source = sources.PlaneSource(plane.source.pulse, "Ex", axis="z", position=0.6 * zc[5] + 0.4 * zc[6])
    return {"pump_40x36x64": pump, "small_soft_24": small, "small_plane_24": plane}


def _compare(gpu, ref, name, bitwise):
    for f in FIELDS:
        a, b = gpu.field(f), ref.field(f)
        if bitwise:
            assert np.array_equal(a, b), (name, f, int(np.count_nonzero(a != b)))
        else:
            assert _rel(a, b) < 1e-5, (name, f, _rel(a, b))
    fa, fb = gpu.film_state(), ref.film_state()
    if fb:
        assert fb["Te"].max() > 300.002, name   # the film was driven
        for k in ("Te", "Tl"):
            assert _rel(fa[k], fb[k]) < 1e-6, (name, k)
        for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
            assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, (name, k)


@pytest.mark.parametrize("case", ["pump_40x36x64", "small_soft_24", "small_plane_24"])
@pytest.mark.parametrize("path", ["step", "run", "launchers"])
def test_ext_call_sequences_match_oracle(case, path):
    import driver
    from mphip.state import FDTDState
    spec = _cases()[case]
    print(f"case ext replay {case} path={path}", flush=True)
    gpu = FDTDState(spec, use_eps_index=False)   # the ext passes no material index
    assert gpu.yee.n_eps == 0 and not gpu._st.record_dft
    ref = driver.OracleState(spec)
    if path == "step":
        for _ in range(100):
            gpu.step()                           # step! -> one mp_run_steps call
    elif path == "run":
        gpu.run(100)                             # run! -> one call for every step
    else:
        ext = ExtLaunchers(gpu)
        for _ in range(100):
            ext.step()
    ref.run(100)
    gpu.synchronize()
    assert gpu.n == 100 and abs(gpu.t - ref.t) <= 1e-12 * ref.t
    # the pump's gyration feeds the FP64 film state back into E (last-ULP exp): coupled bar
    _compare(gpu, ref, case, bitwise=(case != "pump_40x36x64"))
    gpu.free()
