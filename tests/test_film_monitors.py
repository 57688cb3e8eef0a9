"""Film averages, switched fraction, switching diagnostics and the relax / equilibration
phase drivers (CPU: oracle reductions and host algebra).

The oracle's sums are pinned against numpy on random films and against a hand-counted
four-cell case; the host algebra (means, fractions, the reference's masked means and
mixture check) against the definitions of switching_metrics (Monitors.jl:682-727); the
relax phase against run_phase!(::Relax) semantics (Experiment.jl:394-417): relax_step
leaves n and t, the phase advances them by one relax step each and closes the window.
"""
import math

import numpy as np
import pytest

FILM_KEYS = ("m_TM_x", "m_TM_y", "m_TM_z", "m_RE_x", "m_RE_y", "m_RE_z", "Te", "Tl", "Ts_TM", "Ts_RE")


def _oracle(dtype="f64"):
    import driver
    from mphip import configs
    spec, _ = configs.small_coupled(dtype=dtype, steps=10, n=24)
    return driver.OracleState(spec)


def _random_film(ref, rng):
    n = int(ref.film.n)
    vals = {k: rng.uniform(-1, 1, n) for k in FILM_KEYS[:6]}
    for k in FILM_KEYS[6:]:
        vals[k] = rng.uniform(300, 2000, n)
    vals["U_abs"] = rng.uniform(0, 1e3, n)
    vals["U_abs"][[3, n // 2, n - 1]] = 2e3       # three equal maxima: the first wins
    for k, v in vals.items():
        ref.set_film(k, v)
    return {k: ref.fa[k].astype(np.float64) for k in vals}


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_oracle_film_average_matches_numpy(dtype):
    import driver
    ref = _oracle(dtype)
    v = _random_film(ref, np.random.default_rng(1))
    s = driver.film_average(ref)
    want = [v[k].sum() for k in FILM_KEYS[:6]]
    want += [np.sqrt(v["m_TM_x"] ** 2 + v["m_TM_y"] ** 2 + v["m_TM_z"] ** 2).sum(),
             np.sqrt(v["m_RE_x"] ** 2 + v["m_RE_y"] ** 2 + v["m_RE_z"] ** 2).sum()]
    want += [v[k].sum() for k in FILM_KEYS[6:]] + [float(np.sum(v["m_TM_x"] < 0))]
    np.testing.assert_allclose(s, want, rtol=1e-12, atol=1e-9)
    assert s[12] == float(np.sum(v["m_TM_x"] < 0))


def _switching_numpy(mtm, mre, U, core_fraction):
    """switching_metrics restated with numpy masks (Monitors.jl:694-727)."""
    n = len(mtm)
    sw = (mtm < 0) & (mre > 0)
    core = U >= core_fraction * U.max()
    mean = lambda v, m: v[m].sum() / m.sum() if m.sum() else math.nan
    f = sw.sum() / n
    return dict(final_switch_fraction=f, final_tm_reversed_fraction=np.sum(mtm < 0) / n,
                final_re_reversed_fraction=np.sum(mre > 0) / n,
                final_switched_cell_mean_m_TM_x=mean(mtm, sw), final_switched_cell_mean_m_RE_x=mean(mre, sw),
                final_unswitched_cell_mean_m_TM_x=mean(mtm, ~sw), final_unswitched_cell_mean_m_RE_x=mean(mre, ~sw),
                final_mixture_check_m_TM_x=f * mean(mtm, sw) + (1 - f) * mean(mtm, ~sw),
                final_core_cell_count=int(core.sum()),
                final_core_switch_fraction=(sw & core).sum() / core.sum(),
                final_core_mean_m_TM_x=mean(mtm, core), final_core_mean_m_RE_x=mean(mre, core),
                hot_cell_index=int(np.argmax(U)))


@pytest.mark.parametrize("core_fraction", [0.5, 0.9, 0.0])
def test_oracle_switching_metrics_match_numpy(core_fraction):
    import driver
    from mphip.monitors import switching_from_sums
    ref = _oracle()
    v = _random_film(ref, np.random.default_rng(2))
    n = int(ref.film.n)
    got = switching_from_sums(driver.switching_sums(ref, core_fraction), n, core_fraction)
    want = _switching_numpy(v["m_TM_x"], v["m_RE_x"], v["U_abs"], core_fraction)
    for k, w in want.items():
        assert got[k] == pytest.approx(w, rel=1e-12, abs=1e-12), k
    assert got["hot_cell_index"] == 3
    assert got["core_Uabs_fraction_of_max"] == core_fraction


def test_switching_metrics_hand_count():
    """Four cells, counted by hand."""
    from mphip.monitors import switching_from_sums
    import driver
    ref = _oracle()
    n = int(ref.film.n)
    mtm = np.full(n, 0.5)
    mre = np.full(n, -0.5)
    U = np.zeros(n)
    mtm[:4] = [-0.5, 0.3, -0.2, 0.9]
    mre[:4] = [0.4, -0.1, -0.3, 0.2]
    U[:4] = [1.0, 3.0, 3.0, 0.5]
    ref.set_film("m_TM_x", mtm)
    ref.set_film("m_RE_x", mre)
    ref.set_film("U_abs", U)
    s = driver.switching_sums(ref, 0.5)
    assert list(s[:3]) == [1.0, 2.0, 2.0]          # switched: cell 0; TM < 0: 0, 2; RE > 0: 0, 3
    assert s[3] == -0.5 and s[4] == 0.4
    assert s[7] == 3.0 and s[8] == 1.0              # first of the two maxima
    assert s[9] == 2.0 and s[10] == 0.0             # core U >= 1.5: cells 1, 2, none switched
    assert s[11] == pytest.approx(0.1) and s[12] == pytest.approx(-0.4)
    m = switching_from_sums(s, n, 0.5)
    assert m["final_switch_fraction"] == 1 / n
    assert m["final_core_switch_fraction"] == 0.0 and m["hot_cell_index"] == 1
    assert m["final_mixture_check_m_TM_x"] == pytest.approx(float(np.mean(mtm)), rel=1e-12)


def test_empty_film_metrics_are_nan():
    from mphip.monitors import film_values_from_sums, switching_from_sums
    m = switching_from_sums(None, 0)
    assert math.isnan(m["final_switch_fraction"]) and m["final_core_cell_count"] == 0
    v = film_values_from_sums(np.zeros(13), 0)
    assert math.isnan(v["Te_avg"]) and math.isnan(v["mag_TM_norm_avg"])


def test_film_values_from_sums_keys_and_moments():
    from mphip.monitors import FILM_AVERAGE_KEYS, film_values_from_sums
    # the 27 keys of _FILM_AVERAGE_KEYS (Monitors.jl:141-151), in order
    assert len(FILM_AVERAGE_KEYS) == 27 and FILM_AVERAGE_KEYS[0] == "mag_time"
    s = np.arange(13, dtype=float)
    v = film_values_from_sums(s, 4, Ms_TM=2.0, Ms_RE=3.0)
    assert set(v) == set(FILM_AVERAGE_KEYS[1:])
    assert v["m_TM_y_reduced"] == 0.25 and v["M_RE_z_Apm"] == 3.0 * 5 / 4
    assert v["M_net_x_Apm"] == 2.0 * 0 / 4 + 3.0 * 3 / 4
    assert v["Ts_avg"] == 0.5 * (10 / 4 + 11 / 4) and v["mag_RE_norm_avg"] == 7 / 4


def test_phase_steps():
    from mphip.phases import phase_steps
    assert phase_steps(5, 1.0, 1e-15) == 5
    assert phase_steps(None, 10.5e-15, 1e-15) == 11
    assert phase_steps(None, None, 1e-15, 7) == 7
    assert phase_steps(-3, None, 1e-15) == 0


class _HostRecorder:
    def __init__(self, every):
        self.every, self.seen = every, []

    def due(self, n):
        return n % self.every == 0

    def record(self, st):
        self.seen.append((st.n, st.t, float(st.film_state()["Te"].mean())))


def test_relax_phase_and_equilibration_semantics():
    from mphip.phases import equilibrate, run_relax
    ref = _oracle()
    ref.run(4)
    n0, t0 = ref.n, ref.t
    ref.relax_step(1e-15)
    assert (ref.n, ref.t) == (n0, t0)              # relax_step! leaves n and t
    equilibrate(ref, steps=3, dt=0.5e-15)
    assert (ref.n, ref.t) == (n0, t0)
    rec = _HostRecorder(every=2)
    info = run_relax(ref, steps=6, dt=2e-15, subcycles=2, monitors=[rec])
    assert ref.n == n0 + 6 and ref.t == pytest.approx(t0 + 6 * 2e-15, rel=1e-14)
    assert ref.last_mp_n == ref.n
    assert [s[0] for s in rec.seen] == [n for n in range(n0 + 1, n0 + 7) if n % 2 == 0]
    assert info["steps"] == 6 and info["dt"] == 2e-15
    info = run_relax(ref, until=5e-15, dt=2e-15)
    assert info["steps"] == 3
