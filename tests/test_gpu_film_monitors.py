"""Film averages, switched fraction, hot-cell trace and switching diagnostics on the GPU.

The device sums (mp_film_average, mp_switching_metrics) are compared with the oracle on
identical random films: counts, the maximum and its first index exactly, sums within
1e-12 of the sum of absolute terms (FP64 reassociation).  The series monitors are then
driven through a pump phase and a relax phase and compared record by record with numpy
on the film the device holds at each record.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FILM_KEYS = ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z", "Te", "Tl", "Ts_TM", "Ts_RE")


def _spec(dtype):
    from mphip import configs
    spec, _ = configs.coupled_pump(nx=96, ny=96, nz=64, steps=200, fluence_mj_cm2=20.0, npml=6, src_z=40e-9,
                                   n_planes=0, dtype=dtype)
    return spec


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_film_reductions_match_oracle(dtype):
    import driver
    from mphip.monitors import film_average, switching_metrics, switching_from_sums, film_values_from_sums
    from mphip.state import FDTDState
    spec = _spec(dtype)
    gpu = FDTDState(spec)
    ref = driver.OracleState(spec)
    n = int(gpu.film.n)
    print(f"case film reductions {dtype} n={n}", flush=True)
    assert n > 256 * 256          # several grid-stride passes per thread
    rng = np.random.default_rng(5)
    vals = {k: rng.uniform(-1, 1, n) for k in FILM_KEYS[:6]}
    for k in FILM_KEYS[6:]:
        vals[k] = rng.uniform(300, 2000, n)
    U = rng.uniform(0, 1e3, n)
    ties = [70001, 5, 40000, n - 1]          # equal maxima in different workgroups
    U[ties] = 5e3
    vals["U_abs"] = U
    for k, v in vals.items():
        gpu.set_film(k, v)
        ref.set_film(k, v)
    from mphip import _lib
    from mphip.state import DeviceBuffer
    import ctypes as C
    out = DeviceBuffer.zeros(8 * 13)
    _lib.call("film_average", C.byref(gpu.film), out.vp)
    got = out.download(np.float64)
    want = driver.film_average(ref)
    h = {k: ref.fa[k].astype(np.float64) for k in FILM_KEYS}
    scale = [np.abs(h[k]).sum() for k in FILM_KEYS[:6]] + [want[6], want[7]] + \
            [np.abs(h[k]).sum() for k in FILM_KEYS[6:]] + [1.0]
    assert np.all(np.abs(got - want) <= 1e-12 * np.asarray(scale)), np.abs(got - want)
    assert got[12] == want[12]
    for cf in (0.5, 0.999, 0.0):
        _lib.call("switching_metrics", C.byref(gpu.film), cf, out.vp)
        g = out.download(np.float64)
        w = driver.switching_sums(ref, cf)
        for i in (0, 1, 2, 7, 8, 9, 10):
            assert g[i] == w[i], (cf, i, g[i], w[i])
        assert g[8] == 5.0                       # the first of the equal maxima
        for i in (3, 4, 5, 6, 11, 12):
            assert abs(g[i] - w[i]) <= 1e-12 * n, (cf, i)
    out.free()
    m = switching_metrics(gpu, 0.5)
    assert m == pytest.approx(switching_from_sums(driver.switching_sums(ref, 0.5), n, 0.5), rel=1e-12)
    v = film_average(gpu)
    gd = spec.gd
    # M_net = Ms_TM * <m_TM> + Ms_RE * <m_RE> cancels: compare on the scale of its terms
    Ms = float(gd["Ms_TM"]) + float(gd["Ms_RE"])
    assert v == pytest.approx(film_values_from_sums(want, n, float(gd["Ms_TM"]), float(gd["Ms_RE"])), rel=1e-12,
                              abs=1e-12 * Ms)
    gpu.free()


class _HostFilm:
    """Downloads the film at each record (the reference's host view)."""

    def __init__(self, every):
        self.every, self.rows = every, []

    def due(self, n):
        return n % self.every == 0

    def record(self, st):
        f = st.film_state()
        self.rows.append((st.n, st.t, {k: f[k].astype(np.float64) for k in FILM_KEYS + ("U_abs",)}))


def test_series_monitors_through_pump_and_relax():
    from mphip.monitors import FilmAverage, HotCellTrace, SwitchedFraction, switching_metrics
    from mphip.phases import equilibrate, run_pump, run_relax
    from mphip.state import FDTDState
    spec = _spec("f32")
    gpu = FDTDState(spec)
    print("case film series through pump (120) and relax (60)", flush=True)
    equilibrate(gpu, steps=20, dt=0.5e-15)
    assert gpu.n == 0 and gpu.t == 0.0
    avg, swf, host = FilmAverage(every=20), SwitchedFraction(every=20), _HostFilm(every=20)
    hot = HotCellTrace(cell=int(gpu.film.n) // 2, every=20)
    mons = [avg, swf, hot, host]
    info = run_pump(gpu, steps=120, monitors=mons)
    assert info["n_end"] == 120 and gpu.n == 120
    gpu.flush_multiphysics()
    t_pump = gpu.t
    info = run_relax(gpu, steps=60, dt=5e-15, subcycles=4, monitors=mons)
    assert gpu.n == 180 and gpu.t == pytest.approx(t_pump + 60 * 5e-15, rel=1e-12)
    assert gpu.last_mp_n == 180
    d = avg.data()
    s = swf.data()
    hd = hot.data()
    assert len(host.rows) == 9 and len(d["mag_time"]) == 9 and len(s["values"]) == 9
    Ms_TM, Ms_RE = float(spec.gd["Ms_TM"]), float(spec.gd["Ms_RE"])
    for r, (n, t, f) in enumerate(host.rows):
        assert d["mag_time"][r] == t and s["t"][r] == t
        for key, want in (("m_TM_x_reduced", f["m_TM_x"].mean()), ("m_RE_z_reduced", f["m_RE_z"].mean()),
                          ("Te_K", f["Te"].mean()), ("Ts_RE_K", f["Ts_RE"].mean()),
                          ("mag_TM_norm_avg", np.sqrt(f["m_TM_x"] ** 2 + f["m_TM_y"] ** 2 + f["m_TM_z"] ** 2).mean()),
                          ("M_net_x_Apm", Ms_TM * f["m_TM_x"].mean() + Ms_RE * f["m_RE_x"].mean())):
            assert d[key][r] == pytest.approx(want, rel=1e-12, abs=1e-12), (r, key)
        assert s["values"][r] == np.mean(f["m_TM_x"] < 0)
        c = hot.cell
        assert (hd["Te_K"][r], hd["mx_TM"][r], hd["mx_RE"][r]) == (f["Te"][c], f["m_TM_x"][c], f["m_RE_x"][c])
    assert d["Te_K"].max() > 300.5           # the pump heated the film
    f = host.rows[-1][2]
    m = switching_metrics(gpu, 0.5)
    sw = (f["m_TM_x"] < 0) & (f["m_RE_x"] > 0)
    assert m["final_switch_fraction"] == sw.mean()
    assert m["hot_cell_index"] == int(np.argmax(f["U_abs"]))
    core = f["U_abs"] >= 0.5 * f["U_abs"].max()
    assert m["final_core_cell_count"] == int(core.sum())
    assert m["final_core_mean_m_TM_x"] == pytest.approx(f["m_TM_x"][core].mean(), rel=1e-12)
    for mon in (avg, swf):
        mon.free()
    gpu.free()
