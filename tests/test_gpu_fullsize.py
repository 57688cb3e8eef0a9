"""Parity at the full workload sizes (GPU against the oracle, a few steps).

The 384^3 EM stencil (64 random dielectric boxes, a Drude-Lorentz sphere of radius 40
cells, CPML 16, plane source), the 256 x 256 x 192 and the 512 x 512 x 256 coupled pump
(graded z axis, film with 4TM + LLB, gyration, 6 DFT planes x 4 bins).  Both sides start
from the same seeded random E, H and D volumes, so every cell, every CPML slab and every
dispersive list entry does real arithmetic from the first step.  Fields must agree bit for
bit; film temperatures to 1e-6 relative L2 and magnetisation to 1e-4 (FP64 exp differs in
the last ULP between the two math libraries); DFT accumulators to 1e-12 relative L2.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("ex", "ey", "ez", "hx", "hy", "hz", "dx", "dy", "dz")


def _seed_fields(gpu, ref, shape, seed):
    rng = np.random.default_rng(seed)
    for name, scale in (("ex", 1e3), ("ey", 1e3), ("ez", 1e3), ("hx", 3.0), ("hy", 3.0), ("hz", 3.0)):
        v = (rng.standard_normal(shape, dtype=np.float32) * np.float32(scale))
        gpu.set_field(name, v)
        ref.set_field(name, v)
    for name, e in (("dx", "ex"), ("dy", "ey"), ("dz", "ez")):
        v = ref.field(e).astype(np.float64) * 8.854187817e-12
        gpu.set_field(name, v.astype(np.float32))
        ref.set_field(name, v.astype(np.float32))


def _compare(gpu, ref, film=True):
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.array_equal(a, b), (name, int(np.count_nonzero(a != b)))
    if film:
        fa, fb = gpu.film_state(), ref.film_state()
        for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
            d = np.linalg.norm(fa[k].astype(np.float64) - fb[k]) / np.linalg.norm(fb[k])
            assert d < 1e-6, k
        for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
            assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k


@pytest.mark.timeout(900)
def test_em_stencil_384_full_size():
    import driver
    from mphip import configs
    from mphip.state import FDTDState
    spec, meta = configs.em_stencil(n=384, npml=16, steps=5000)
    print(f"case em_stencil 384^3 ({meta['n_ade']} ADE cells), 3 steps", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    _seed_fields(gpu, ref, spec.shape, 1)
    gpu.run(3)
    ref.run(3)
    gpu.synchronize()
    _compare(gpu, ref, film=False)
    for c in range(3):
        g, r = gpu.ade_state(c), ref.ade_arr[c]
        assert np.array_equal(g["P"], r["P"]) and np.array_equal(g["J"], r["J"]), c
    gpu.free()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("shape", [(256, 256, 192), (512, 512, 256)])
def test_coupled_pump_full_size(shape):
    import driver
    from mphip import configs
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    from test_gpu_parity import OracleDFT, rel_l2
    nx, ny, nz = shape
    spec, meta = configs.coupled_pump(nx=nx, ny=ny, nz=nz, steps=20000)
    print(f"case coupled_pump {nx}x{ny}x{nz} ({spec.n_film} film cells), 3 steps", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    _seed_fields(gpu, ref, spec.shape, 2)
    mon = DFTPlanes(gpu, meta["planes"], meta["omega_bins"])
    gpu.attach_dft(mon)
    ref_mon = OracleDFT(ref, meta["planes"], meta["omega_bins"])
    gpu.run(3)
    for _ in range(3):
        ref.step()
        ref_mon.record(ref)
    gpu.synchronize()
    _compare(gpu, ref)
    for p in range(len(meta["planes"])):
        g, r = mon.download(p), ref_mon.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            assert rel_l2(g[k], r[k]) < 1e-12, (p, k)
    gpu.free()


@pytest.mark.timeout(900)
def test_coupled_pump_bf16_storage_full_size():
    """The 512 x 512 x 256 pump with CPML psi and ADE P/J stored in BF16 (the reduced
    storage form), 3 steps from seeded fields: fields, raw psi and raw P/J bit for bit
    against the oracle's restatement of the same storage."""
    import driver
    from mphip import configs
    from mphip.state import FDTDState
    spec, meta = configs.coupled_pump(nx=512, ny=512, nz=256, steps=20000, psi_precision="bf16",
                                      ade_state_precision="bf16")
    print(f"case coupled_pump 512x512x256 bf16 storage ({spec.n_film} film cells), 3 steps", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    _seed_fields(gpu, ref, spec.shape, 4)
    gpu.run(3)
    ref.run(3)
    gpu.synchronize()
    _compare(gpu, ref)
    for s in range(12):
        assert np.array_equal(gpu.psi("h", s), ref.psi_h[s]), ("psi_h", s)
        assert np.array_equal(gpu.psi("d", s), ref.psi_d[s]), ("psi_d", s)
    for c in range(3):
        g, r = gpu.ade_state(c), ref.ade_arr[c]
        assert np.array_equal(g["P"], r["P"]) and np.array_equal(g["J"], r["J"]), c
    gpu.free()


@pytest.mark.timeout(900)
def test_surrogate_step_full_film():
    """The learned film update on the 2.62 M-cell film of the 512 x 512 x 256 pump:
    mp_surrogate_step on every cell against orc_surrogate_apply on every 61st cell, from
    seeded film states, with all steps accepted (tol 1e9) and with the midpoint check at
    1e-4 (a mix of accepted and fallback cells)."""
    import ctypes as C
    import driver
    from mphip import _lib, abi, configs
    from mphip.state import DeviceBuffer, FDTDState
    from mphip.surrogate import film_rows
    from test_gpu_surrogate import _net
    spec, _ = configs.coupled_pump(nx=512, ny=512, nz=256, steps=20000)
    N = spec.n_film
    print(f"case surrogate step on the full film ({N} cells)", flush=True)
    L = driver.lib(abi)
    sur, _, _ = _net(8)
    rng = np.random.default_rng(21)
    vals = dict(Te=rng.uniform(300, 2500, N), Tl=rng.uniform(300, 700, N), Ts_TM=rng.uniform(300, 900, N),
                Ts_RE=rng.uniform(300, 900, N), m_TM_x=rng.uniform(-0.9, 0.9, N), m_RE_x=rng.uniform(-0.9, 0.9, N),
                window_sum=rng.uniform(0, 1e22, N) * spec.dt)
    sample = np.arange(0, N, 61)
    n = len(sample)
    for tol in (1e9, 1e-4):
        gpu = FDTDState(spec)
        for k, v in vals.items():
            gpu.set_film(k, v)
        init = gpu.film_state()
        X, _ = film_rows(gpu, gpu.dt, 4, True)
        rates = sur.forward(DeviceBuffer.from_array(X), N).download(np.float32).reshape(-1, 6)[:N]
        sur.reset_fallbacks()
        _lib.call("surrogate_step", C.byref(gpu.film), gpu.gd_arr.ctypes.data_as(C.c_void_p), C.byref(gpu.lut),
                  C.byref(sur.mlp), gpu.dt, 4, spec.brillouin_iters, 1, tol, sur.fallbacks.vp)
        got = gpu.film_state()
        # the oracle on the sampled cells alone (every cell is updated on its own)
        fa = {k: np.ascontiguousarray(init[k][sample]) for k in driver.OracleState.FILM_NAMES}
        fa["window_sum"] = np.ascontiguousarray(init["window_sum"][sample].astype(np.float64))
        for k in ("U_abs", "peak"):
            fa[k] = np.zeros(n)
        fa["cells"] = np.arange(n, dtype=np.int32)
        for c in "xyz":
            fa["pos_" + c] = np.full(n, -1, np.int32)
        fm = abi.Film()
        fm.n, fm.dtype = n, abi.MP_F32
        for k, a in fa.items():
            setattr(fm, k, a.ctypes.data)
        tm, re_, T_min, inv_dT, nT = spec.lut
        lut_tm, lut_re = tm.copy(), re_.copy()
        lut = abi.Lut()
        lut.tm, lut.re, lut.T_min, lut.inv_dT, lut.n, lut.dtype = (lut_tm.ctypes.data, lut_re.ctypes.data, T_min,
                                                                  inv_dT, nT, abi.MP_F32)
        gd = np.ascontiguousarray(spec.gd.as_array(np.float64))
        r_rates = np.ascontiguousarray(rates[sample])
        fb = C.c_int64(0)
        assert L.orc_surrogate_apply(C.byref(fm), gd.ctypes.data, C.byref(lut), r_rates.ctypes.data, spec.dt, 4,
                                     spec.brillouin_iters, 1, tol, C.byref(fb)) == 0
        gf = sur.fallback_count()
        print(f"tol {tol}: fallbacks gpu {gf} of {N} ({gf / N:.3f}), oracle {fb.value} of {n} "
              f"({fb.value / n:.3f})", flush=True)
        assert abs(gf / N - fb.value / n) < 0.02
        for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
            assert np.median(np.abs(got[k][sample] / fa[k] - 1)) < 1e-6, k
        for k in ("m_TM_x", "m_RE_x", "m_TM_y", "m_RE_z"):
            assert np.median(np.abs(got[k][sample].astype(np.float64) - fa[k])) < 1e-6, k
        gpu.free()
    sur.free()


@pytest.mark.timeout(900)
def test_coupled_pump_workload_from_rest():
    """The 512 x 512 x 256 pump workload itself, not seeded fields: from rest with the pump
    pulse already on (t0 is half the 20 000-step run, inside the pulse), 40 steps of the
    step driver with its batched DFT record on the 6 planes x 4 bins, against the oracle
    stepping and recording every step.  Fields within 1e-5 relative L2 (the gyration feeds
    the FP64 film state back into E), film to the film bar, DFT accumulators to 1e-10
    relative L2 (cos/sin of the two math libraries)."""
    import driver
    from mphip import configs
    from mphip.monitors import DFTPlanes
    from mphip.state import FDTDState
    spec, meta = configs.coupled_pump(nx=512, ny=512, nz=256, steps=20000)
    print("case coupled pump 512x512x256 from rest, 40 steps, DFT on", flush=True)
    gpu, ref = FDTDState(spec), driver.OracleState(spec)
    mon = DFTPlanes(gpu, meta["planes"], meta["omega_bins"])
    gpu.attach_dft(mon)
    ref_mon = driver.OracleDFT(ref, meta["planes"], meta["omega_bins"])
    gpu.run(40)
    for _ in range(40):
        ref.step()
        ref_mon.record(ref)
    gpu.synchronize()
    for name in FIELDS:
        a, b = gpu.field(name), ref.field(name)
        assert np.linalg.norm(b) > 0, name
        d = np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64))
        assert d < 1e-5, (name, d, int(np.count_nonzero(a != b)))
    fa, fb = gpu.film_state(), ref.film_state()
    assert fb["U_abs"].max() > 0
    for k in ("Te", "Tl", "Ts_TM", "Ts_RE"):
        d = np.linalg.norm(fa[k].astype(np.float64) - fb[k]) / np.linalg.norm(fb[k])
        assert d < 1e-6, k
    for k in ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z"):
        assert np.abs(fa[k].astype(np.float64) - fb[k]).max() < 1e-4, k
    reached = 0
    for p in range(len(meta["planes"])):
        g, r = mon.download(p), ref_mon.download(p)
        for k in ("ea", "eb", "ha", "hb", "energy"):
            den = np.linalg.norm(r[k])
            if den == 0:    # the z plane 45 cells below the source is not reached in 40 steps
                assert not np.any(g[k]), (p, k)
                continue
            reached += 1
            assert np.linalg.norm(g[k] - r[k]) / den < 1e-10, (p, k)
    assert reached >= 20
    mon.free()
    gpu.free()
