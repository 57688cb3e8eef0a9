"""Probe read-out, polarimetry and transmission: host algebra and the oracle's plane
reductions (CPU).

Known answers are the reference's own (test/runtests.jl:517-579: the analytic forward
plane wave through a ProbeReadout, reference normalisation, MO consistency metrics); the
oracle's reductions are pinned against numpy's complex arithmetic on random planes and
against closed forms of the plane wave (real, positive Sx equal to |Ez(w)|^2 A / Z0).
"""
import ctypes as C
import math

import numpy as np
import pytest

Z0 = 376.730313668


def _uniform_spec(nx=7, ny=4, nz=4, dx=1e-7, dyz=5e-8, precision="f64"):
    from mphip import geometry, grid
    from mphip.spec import build_state_spec
    g = grid.Grid3D(grid.uniform_axis_n(nx, dx), grid.uniform_axis_n(ny, dyz, -1e-7),
                    grid.uniform_axis_n(nz, dyz, -1e-7))
    geo = geometry.rasterize(geometry.Scene(), g)
    return build_state_spec(g, geo, 1e-17, compute_precision=precision, field_precision=precision)


# ---- host algebra --------------------------------------------------------------------------

def test_jones_angles_closed_forms():
    from mphip.probe import probe_jones_angles_deg as ang
    for phi in (0.0, 10.0, -30.0, 44.0):
        r = math.radians(phi)
        a = ang(math.sin(r), math.cos(r))          # linear, phi from the Ez axis
        assert a["rotation_deg"] == pytest.approx(phi, abs=1e-12)
        assert a["ellipticity_deg"] == pytest.approx(0.0, abs=1e-12)
    a = ang(-1j, 1.0)                               # Ey lags Ez by 90 deg: circular
    assert a["ellipticity_deg"] == pytest.approx(-45.0, abs=1e-12)
    assert ang(1j, 1.0)["ellipticity_deg"] == pytest.approx(45.0, abs=1e-12)
    assert ang(1.0, 0.0)["rotation_deg"] == pytest.approx(90.0, abs=1e-12)
    # overall phase and scale do not change the angles
    b = ang(0.3 * np.exp(0.7j) * (0.2 + 0.5j), 0.3 * np.exp(0.7j) * (1.1 - 0.1j))
    c = ang(0.2 + 0.5j, 1.1 - 0.1j)
    assert b["rotation_deg"] == pytest.approx(c["rotation_deg"], abs=1e-12)
    assert b["ellipticity_deg"] == pytest.approx(c["ellipticity_deg"], abs=1e-12)
    n = ang(0.0, 0.0)
    assert math.isnan(n["rotation_deg"]) and math.isnan(n["ellipticity_deg"])
    assert math.isnan(ang(complex(math.inf, 0), 1.0)["rotation_deg"])


def test_mo_consistency_metrics_known_answers():
    """runtests.jl:567-579."""
    from mphip.probe import mo_consistency_metrics
    plus = dict(T_plus_R_plus_A=0.999, theta_faraday_deg=2.0)
    zero = dict(T_plus_R_plus_A=1.000, theta_faraday_deg=0.0)
    minus = dict(T_plus_R_plus_A=1.001, theta_faraday_deg=-1.0)
    m = mo_consistency_metrics(plus, zero, minus)
    assert m["even_in_M_violation_plus_minus"] == pytest.approx(-0.002, abs=1e-12)
    assert m["TRA_spread_plus_zero_minus"] == pytest.approx(0.002, abs=1e-12)
    assert m["pedestal_avg_vs_zeroM"] == [0.5, 0.0]
    assert m["conservative"]
    assert not mo_consistency_metrics(plus, zero, dict(minus, T_plus_R_plus_A=1.010))["conservative"]
    with pytest.raises(ValueError):
        mo_consistency_metrics(plus, zero, dict(minus, T_plus_R_plus_A=math.nan))
    with pytest.raises(ValueError):
        mo_consistency_metrics(plus, zero, dict(minus, theta_faraday_deg=math.inf))
    # totals from T, R, A when T_plus_R_plus_A is absent
    m2 = mo_consistency_metrics(dict(T=0.5, R=0.25, A=0.25, theta_faraday_deg=1.0), zero, minus)
    assert m2["TRA_plus_M"] == 1.0


def test_energy_balance_and_contrast():
    from mphip.probe import energy_balance, probe_contrast
    e = energy_balance(incident=4.0, reflected=1.0, transmitted=2.0)
    assert (e["T"], e["R"], e["A"]) == (0.5, 0.25, 0.25)
    keys = ("T", "R", "A", "theta_faraday_deg", "theta_kerr_deg", "ellipticity_faraday_deg",
            "ellipticity_kerr_deg", "group_delay_fs", "pulse_broadening_fs")
    a = {k: float(i + 1) for i, k in enumerate(keys)}
    b = {k: 0.5 * v for k, v in a.items()}
    a["T_plus_R_plus_A"], b["T_plus_R_plus_A"] = 6.0, 3.0
    d = probe_contrast(a, b)
    assert d["delta_T"] == 0.5 and d["delta_pulse_broadening_fs"] == 4.5   # initial - switched
    assert d["initial_T_R_A"] == [1.0, 2.0, 3.0] and d["switched_T_plus_R_plus_A"] == 3.0


def test_trace_moments_gaussian():
    from mphip.probe import trace_moments
    t = np.arange(4000) * 0.05e-15
    t0, s = 90e-15, 12e-15
    m = trace_moments(t, np.exp(-((t - t0) ** 2) / (2 * s * s)))
    # |y|^2 = exp(-(t-t0)^2 / s^2): mean t0, rms s / sqrt(2)
    assert m["delay_fs"] == pytest.approx(90.0, rel=1e-9)
    assert m["rms_fs"] == pytest.approx(12.0 / math.sqrt(2), rel=1e-6)
    assert math.isnan(trace_moments([], [])["delay_fs"])
    assert math.isnan(trace_moments([1.0, 2.0], [0.0, 0.0])["rms_fs"])


def test_default_bins_and_centre():
    from mphip.probe import ProbeReadout, default_bins
    bins, w0 = default_bins(532e-9, 5)
    assert len(bins) == 5 and bins[2] == pytest.approx(w0, rel=1e-15)
    assert bins[0] == pytest.approx(0.96 * w0, rel=1e-14) and bins[-1] == pytest.approx(1.04 * w0, rel=1e-14)
    ro = ProbeReadout(pre=1e-7, post=4e-7, spectrum_bins=4)
    assert ro.center_bin == 1          # first of the two nearest, as argmin
    assert ProbeReadout(pre=1e-7, post=4e-7, omega_bins=[1.0, 2.0, w0]).center_bin == 2
    with pytest.raises(ValueError):
        default_bins(532e-9, 1)


def _frame_schedule(ro, first, last):
    """The steps at which the read-out takes a movie frame, with store_frames off (no device)."""
    class _State:
        pass
    st = _State()
    for n in range(first, last + 1):
        st.n, st.t = n, n * 1e-15
        ro._record_frame(st)
    return list(ro.frame_indices)


def test_probe_frame_schedule():
    """_probe_frame_due (ProbeReadout.jl:183-191), schedules worked out by hand from its rules:
    a plan is followed entry by entry; else every frame_skip-th step, or with frame_skip 0
    every ceil(n / frame_target)-th step, until frame_target frames are taken."""
    from mphip.probe import ProbeReadout

    def ro(**kw):
        return ProbeReadout(pre=1e-7, post=4e-7, store_frames=False, **kw)

    assert _frame_schedule(ro(), 1, 50) == []                                   # frame_target 0: none
    assert _frame_schedule(ro(frame_target=3), 1, 50) == [1, 2, 3]              # ceil(n/3) = 1 for n <= 3
    assert _frame_schedule(ro(frame_target=4, frame_skip=5), 1, 50) == [5, 10, 15, 20]
    # from step 101 with frame_target 4: ceil(104/4) = 26 divides 104, ceil(108/4) = 27 divides
    # 108, then 112 (28) and 116 (29); 101-103, 105-107, ... are not multiples of theirs
    assert _frame_schedule(ro(frame_target=4), 101, 200) == [104, 108, 112, 116]
    r = ro(frame_target=3, frame_indices=[3, 7, 11])
    assert r.frame_skip == 4                                                     # plan[2] - plan[1]
    assert _frame_schedule(r, 1, 50) == [3, 7, 11] and r.frame_ptr == 3
    assert _frame_schedule(ro(frame_target=0, frame_indices=[3, 7]), 1, 50) == []
    r = ro(frame_target=5, frame_indices=[2, 4])
    assert _frame_schedule(r, 1, 50) == [2, 4]                                   # the plan ends first
    r = ro(frame_target=2)
    _frame_schedule(r, 1, 10)
    assert r.frame_times == [1e-15, 2e-15]
    shot = r.probe_shot()
    assert shot["probe_frame_count"] == 2 and shot["probe_frame_indices"] == [1, 2]
    assert all(v.shape == (0, 0, 0) for v in shot["frames"].values())


def test_polarimetry_and_tr_summaries():
    from mphip.probe import polarimetry_summary, tr_summary
    w = 2 * math.pi * 3e14
    t = np.arange(2000) * 1e-17
    # Ey = sin(wt), Ez = cos(wt): phasors Ey ~ -i N/2, Ez ~ N/2 over whole periods
    per = 2 * math.pi / w
    t = t[t < 20 * per - 1e-19]
    p = polarimetry_summary(t, np.sin(w * t), np.cos(w * t), w)
    assert abs(p["ellipticity_deg"]) == pytest.approx(45.0, abs=0.5)
    p = polarimetry_summary(t, 0.5 * np.cos(w * t), np.cos(w * t), w)
    assert p["rotation_deg"] == pytest.approx(math.degrees(math.atan2(0.5, 1.0)), abs=1e-6)
    assert math.isnan(polarimetry_summary([], [], [], w)["rotation_deg"])
    s = tr_summary([0.0, 2.0, 4.0], [1.0, 2.0, -0.5], False, 5.0)
    assert s["energy"] == 5.0 and s["ratio"] == 1.0
    s = tr_summary([0.0, 2.0, 4.0], [1.0, 2.0, -0.5], True, 0.0)
    assert s["energy"] == 2.0 * 5.25 and s["ratio"] == s["energy"] / np.finfo(float).eps
    assert tr_summary([1.0], [3.0], False, 1.0)["energy"] == 0.0


# ---- oracle reductions ----------------------------------------------------------------------

def _random_planes(ref, rng, nb):
    import driver
    nx = ref.spec.shape[0]
    dft = driver.OracleDFT(ref, [(0, 1), (0, nx - 2)], np.linspace(1.0, 2.0, nb))
    for p in range(2):
        for k, a in dft.raw(p).items():
            a[...] = rng.standard_normal(a.size)
    return dft


@pytest.mark.parametrize("nb", [1, 5])
def test_oracle_shot_reduce_matches_numpy(nb):
    import driver
    from mphip.monitors import plane_area
    ref = driver.OracleState(_uniform_spec(nx=9, ny=5, nz=3))
    rng = np.random.default_rng(nb)
    dft = _random_planes(ref, rng, nb)
    area = plane_area(ref.spec.grid, 0) * rng.uniform(0.5, 1.5, 15)
    mode_w = rng.standard_normal(15)
    q0 = nb // 2
    out = np.zeros(4 * nb + 10)
    rc = ref.L.orc_probe_shot_reduce(C.byref(dft.desc[0]), C.byref(dft.desc[1]), 5, 3, nb, q0,
                                     mode_w.ctypes.data_as(C.c_void_p), area.ctypes.data_as(C.c_void_p),
                                     out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    A = area.reshape((5, 3), order="F")
    W = mode_w.reshape((5, 3), order="F") * A
    for p, name in ((0, "pre"), (1, "post")):
        d = dft.download(p)
        sx = np.array([np.sum((d["ea"][:, :, q] * np.conj(d["hb"][:, :, q]) -
                               d["eb"][:, :, q] * np.conj(d["ha"][:, :, q])) * A) for q in range(nb)])
        got = out[2 * nb * p:2 * nb * (p + 1)]
        np.testing.assert_allclose(got[0::2], sx.real, rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(got[1::2], sx.imag, rtol=1e-13, atol=1e-13)
        assert out[4 * nb + p] == pytest.approx(float(np.sum(d["energy"] * A)), rel=1e-13, abs=1e-13)
    post, pre = dft.download(1), dft.download(0)
    expect = [np.sum(post["ea"][:, :, q0] * W), np.sum(post["eb"][:, :, q0] * W),
              np.sum(pre["ea"][:, :, q0] * W), np.sum(pre["eb"][:, :, q0] * W)]
    j = out[4 * nb + 2:]
    np.testing.assert_allclose(j[0::2], np.real(expect), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(j[1::2], np.imag(expect), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_oracle_plane_overlap_matches_numpy(axis):
    import driver
    ref = driver.OracleState(_uniform_spec(nx=6, ny=5, nz=4))
    rng = np.random.default_rng(axis)
    vols = {}
    for k in ("ex", "ey", "ez", "hx", "hy", "hz"):
        vols[k] = rng.standard_normal((6, 5, 4))
        ref.set_field(k, vols[k])
    idx = 2
    for comp, k in enumerate(("ex", "ey", "ez", "hx", "hy", "hz")):
        v = vols[k]
        # plane order: x -> (j, k), y -> (k, i), z -> (i, j)
        pl = (v[idx, :, :], v[:, idx, :].T, v[:, :, idx])[axis]
        w = rng.standard_normal(pl.size)
        assert driver.plane_overlap(ref, comp, axis, idx, w) == pytest.approx(
            float(np.sum(pl * w.reshape(pl.shape, order="F"))), rel=1e-13, abs=1e-13)
        assert driver.plane_overlap(ref, comp, axis, idx) == pytest.approx(float(np.sum(pl)), rel=1e-13, abs=1e-13)


def _plane_wave_probe(dt=1e-17, steps=200):
    """The analytic forward wave of runtests.jl:517-565 (Ez = cos(w0 t), Hy = -Ez/Z0 on
    every cell) through the oracle's DFT record and reductions."""
    import driver
    from mphip.probe import ProbeReadout
    spec = _uniform_spec()
    ref = driver.OracleState(spec)
    ro = ProbeReadout(pre=1.0e-7, post=4.0e-7, lambda0=532e-9, trace_stride=1, spectrum_bins=5)
    probe = driver.OracleProbe(ref, ro)
    w0 = 2 * math.pi * 299792458.0 / 532e-9
    shape = spec.shape
    for n in range(1, steps + 1):
        t = n * dt
        ref.n, ref.t = n, t
        ref.set_field("ez", np.full(shape, math.cos(w0 * t)))
        ref.set_field("hy", np.full(shape, -math.cos(w0 * t) / Z0))
        probe.record()
    return ref, ro, probe, w0


def test_probe_spectral_poynting_reference_normalisation():
    """runtests.jl:517-565 on the oracle reductions + the library's shot algebra."""
    from mphip.probe import shot_from_reductions, unpack_reductions
    ref, ro, probe, w0 = _plane_wave_probe()
    nb = 5
    red = unpack_reductions(probe.reductions_raw(), nb)
    shot = shot_from_reductions(red, nb, probe.mode_norm)
    assert np.max(np.abs(shot["Sx_pre_net_re"])) > 0
    assert all((not np.isfinite(r)) or abs(r) <= 1e-9 for r in shot["R_omega"])
    nonref = shot_from_reductions(red, nb, probe.mode_norm, incident_energy=2 * shot["incident_energy_J"],
                                  incident_sx=2 * shot["Sx_inc"])
    fin = [r for r in nonref["R_omega"] if np.isfinite(r)]
    assert any(r != 0.0 for r in fin)
    assert max(fin) == pytest.approx(0.5, abs=1e-6)
    # closed forms of the uniform forward wave: Sx(w) = |Ez(w)|^2 * A / Z0 (real and
    # positive), equal on both planes, so T_omega = 1 and T = 1, R = 0
    g = ref.spec.grid
    A_tot = float(np.sum(np.outer(g.y.widths(), g.z.widths())))
    t = np.arange(1, 201) * 1e-17
    for q, w in enumerate(ro.omega_bins):
        ez_w = np.sum(np.cos(w0 * t) * np.exp(-1j * w * t) * 1e-17)
        assert red["Sx_pre"][q].real == pytest.approx(abs(ez_w) ** 2 * A_tot / Z0, rel=1e-10)
        assert abs(red["Sx_pre"][q].imag) <= 1e-12 * red["Sx_pre"][q].real
    np.testing.assert_allclose(shot["T_omega"], 1.0, rtol=1e-12)
    assert shot["T"] == pytest.approx(1.0, rel=1e-12) and shot["R"] == pytest.approx(0.0, abs=1e-12)
    assert shot["net_energy_pre_J"] == pytest.approx(A_tot * float(np.sum(np.cos(w0 * t) ** 2)) * 1e-17 / Z0,
                                                     rel=1e-12)
    # a pure Ez wave: rotation 0, ellipticity 0, and the mode-weighted trace is Ez itself
    assert shot["theta_faraday_deg"] == pytest.approx(0.0, abs=1e-9)
    assert shot["ellipticity_kerr_deg"] == pytest.approx(0.0, abs=1e-9)
    tr = np.array(probe.traces)
    np.testing.assert_allclose(tr[:, 1], np.cos(w0 * t), rtol=1e-12, atol=1e-15)
    assert np.all(tr[:, 0] == 0.0)


def test_shot_without_records_is_zero():
    from mphip.probe import ProbeReadout
    ro = ProbeReadout(pre=1e-7, post=4e-7)
    s = ro.probe_shot()
    assert s["T"] == 0.0 and s["R"] == 0.0 and len(s["T_omega"]) == 5
    assert np.all(np.isnan(s["T_omega"])) and math.isnan(s["theta_faraday_deg"])
