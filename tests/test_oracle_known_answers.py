"""Pins the C oracle to the reference's own known answers (CPU only).

Each check is a test assertion or documented result of the reference
(test/runtests.jl, docs/src/*.md) replayed through the oracle's restated kernels, or a
closed form of the mathematics.  None of them re-derives the oracle's own formulas.
"""
import ctypes as C
import math

import numpy as np
import pytest


def _lib():
    import driver
    from mphip import abi
    return driver.lib(abi)


def _params_arr(gd):
    return np.ascontiguousarray(gd.as_array(np.float64))


# ---- grids and CPML (test/runtests.jl:38-48, 78-95, 637-650) ------------------------------
def test_not_gate_mesh_exact():
    from mphip import grid
    g = grid.not_gate_reference_grid()
    assert g.shape == (4099, 190, 100)
    assert grid.cfl_dt(g, courant=0.99) == 1.1218649561077176e-17


def test_pml_slab_and_slab_shapes():
    from mphip import cpml, grid
    g = grid.Grid3D(*(grid.uniform_axis(0.0, 1.0e-6, 5e-8) for _ in range(3)))
    N = g.x.n
    dt = grid.cfl_dt(g, courant=0.4)
    assert cpml.pml_slab(5, N, 5) == (5, 1)
    assert cpml.pml_slab(6, N, 5) == (0, 0)
    assert cpml.pml_slab(N - 5 + 1, N, 5) == (1, 2)
    assert cpml.pml_slab(N, N, 5) == (5, 2)
    c = cpml.build_cpml(g, 5, dt)
    assert c.widths == (5, 5, 5)
    assert c.slab_shape(0) == (5, N, N) and c.slab_shape(1) == (N, 5, N) and c.slab_shape(2) == (N, N, 5)
    mid = (N + 1) // 2 - 1
    assert 0.0 < c.axes[0].b[0] < 1.0
    assert c.axes[0].b[mid] == 0.0 and c.axes[0].inv_kappa[mid] == 1.0
    kappa, a, b = cpml.build_cpml_profiles(20, 4, 1e-8, 1e-17)
    assert len(kappa) == 20 and b[0] < 1.0 and a[9] == 0.0


def test_create_pole_and_pump_fit():
    """runtests.jl:44-47 (finite poles) and :167-182 (pump pole set reproduces eps at
    800 nm within 5% through the discrete recursion)."""
    from mphip import dispersion, grid, models
    p = grid.FDTDParams()
    assert all(math.isfinite(c) for c in dispersion.create_pole(0.0, 3e14, 1e30, 1e-18, p.eps0))
    gd = models.ferrimagnet_parameters()
    dt = 1e-17
    poles = dispersion.build_pump_poles(dt, p.eps0, gd)
    chi = sum(dispersion.discrete_pole_chi(pl, dt, p.eps0, gd["omega_pump"]) for pl in poles)
    eps_fit = 1.0 + chi
    assert abs(eps_fit.real - gd["eps_real_pump"]) <= 0.05 * abs(gd["eps_real_pump"])
    assert abs(eps_fit.imag - gd["eps_imag_pump"]) <= 0.05 * abs(gd["eps_imag_pump"])


# ---- model parameters, Brillouin, LUT (runtests.jl:50-61) ----------------------------------
def test_brillouin_and_lut(oracle_lib):
    from mphip import models
    assert abs(models.brillouin(1.0, 1e-8)) < 1e-6
    assert abs(oracle_lib.orc_brillouin_fn(1.0, 1e-8)) < 1e-6
    # closed form: B_S(x) -> coth(x) - 1/x style limits; B_1/2(x) = tanh(x)
    for x in (0.3, 1.7, -2.2, 40.0):
        assert abs(models.brillouin(0.5, x) - math.tanh(x)) < 1e-12
        assert abs(oracle_lib.orc_brillouin_fn(0.5, x) - math.tanh(x)) < 1e-12
    gd = models.ferrimagnet_parameters(T_Curie=550.0, Q_voigt_TM=0.025)
    assert gd["T_Curie"] == 550.0
    assert models.optical_coupling(gd, 1.0, -1.0) > 0.0
    tm, re_, Tmin, invdT, nT = models.build_m_eq_lut(gd, T_max=20.0, dT=5.0)
    assert len(tm) == nT and Tmin == 1.0 and invdT == 0.2
    gd0 = models.ferrimagnet_parameters()
    tm, re_, Tmin, invdT, nT = models.build_m_eq_lut(gd0)
    mtm = models.lookup_m_eq_lut(tm, gd0["T0"], Tmin, invdT, nT)
    mre = models.lookup_m_eq_lut(re_, gd0["T0"], Tmin, invdT, nT)
    assert mtm > 0.9 and mre < -0.9          # antiparallel ferrimagnet at 300 K
    tmc = np.ascontiguousarray(tm)
    for T in (1.0, 299.5, 300.0, 539.2, 2000.0, float("nan")):
        a = models.lookup_m_eq_lut(tm, T, Tmin, invdT, nT)
        b = oracle_lib.orc_lut_lookup(tmc.ctypes.data_as(C.c_void_p), T, Tmin, invdT, nT)
        assert a == b or (math.isnan(T) and a == 0.0 == b)


def _llb(oracle_lib, gd_arr, lut, m, te, tsT, tsR, dt, iters=2):
    tm, re_, Tmin, invdT, nT = lut
    mm = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros(4)
    oracle_lib.orc_llb_step(mm.ctypes.data_as(C.c_void_p), te, tsT, tsR, gd_arr.ctypes.data_as(C.c_void_p),
                            tm.ctypes.data_as(C.c_void_p), re_.ctypes.data_as(C.c_void_p), Tmin, invdT, nT,
                            dt, iters, out.ctypes.data_as(C.c_void_p))
    return mm, out


def test_llb_all_optical_switching(oracle_lib):
    """runtests.jl:110-127: 6000 x 0.5 fs at a 1200 K plateau then a ramp reverses both
    sublattices; |m_TM| stays capped at 0.992."""
    from mphip import models
    gd = models.ferrimagnet_parameters()
    tm, re_, Tmin, invdT, nT = models.build_m_eq_lut(gd)
    lut = (np.ascontiguousarray(tm), np.ascontiguousarray(re_), Tmin, invdT, nT)
    mtm = models.lookup_m_eq_lut(tm, gd["T0"], Tmin, invdT, nT)
    mre = models.lookup_m_eq_lut(re_, gd["T0"], Tmin, invdT, nT)
    m = np.array([mtm, 0.0, 1e-4, mre, 0.0, 1e-4])
    garr = _params_arr(gd)
    for n in range(1, 6001):
        T = 1200.0 if n < 1700 else max(300.0, 1200.0 - (n - 1700) * 0.4)
        m, _ = _llb(oracle_lib, garr, lut, m, T, T, T, 0.5e-15)
    assert m[0] < -0.8
    assert m[3] > 0.8
    assert math.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2) <= 0.992 + 1e-9


def test_4tm_relaxes_to_ambient(oracle_lib):
    """runtests.jl:129-139: from 1500 K, 200000 x 1 fs with no heating."""
    from mphip import models
    gd = models.ferrimagnet_parameters()
    garr = _params_arr(gd)
    T4 = np.array([1500.0, 1500.0, 1500.0, 1500.0])
    p = T4.ctypes.data_as(C.c_void_p)
    g = garr.ctypes.data_as(C.c_void_p)
    for _ in range(200000):
        oracle_lib.orc_update_4tm_cell(p, 0.0, 0.0, 0.0, g, 1e-15)
    assert 290.0 < T4[1] < 360.0
    assert np.all(np.isfinite(T4))


def test_4tm_electron_energy_closed_form(oracle_lib):
    """update_Te_energy integrates u_e = gamma_e Te^2 / 2 explicitly: with the baths at
    the electron temperature one step adds exactly pabs*dt to u_e."""
    from mphip import models
    gd = models.ferrimagnet_parameters()
    g = _params_arr(gd)
    T4 = np.array([800.0, 800.0, 800.0, 800.0])
    pabs, dt = 3.0e21, 1e-16
    oracle_lib.orc_update_4tm_cell(T4.ctypes.data_as(C.c_void_p), pabs, 0.0, 0.0, g.ctypes.data_as(C.c_void_p), dt)
    ue0 = 0.5 * gd["gamma_e"] * 800.0 ** 2
    assert T4[0] == pytest.approx(math.sqrt(2.0 * (ue0 + dt * pabs) / gd["gamma_e"]), rel=1e-14)
    # lattice: only the 2 ps relaxation toward T0 acts when Te == Tl
    assert T4[1] == pytest.approx(800.0 - dt * (800.0 - gd["T0"]) / 2e-12, rel=1e-14)


def test_aos_switching_0d_documented(oracle_lib):
    """examples/aos_switching_0d.jl with the result documented in
    docs/src/magneto_optics.md:83-87: FeCo crosses zero at 0.92 ps, Gd at 1.47 ps,
    T_e peaks at 2381 K, final m_x -0.990 / +0.968."""
    from mphip import models
    gd = models.ferrimagnet_parameters()
    tm, re_, Tmin, invdT, nT = models.build_m_eq_lut(gd)
    lut = (np.ascontiguousarray(tm), np.ascontiguousarray(re_), Tmin, invdT, nT)
    m = np.array([models.lookup_m_eq_lut(tm, gd["T0"], Tmin, invdT, nT), 0.0, 1e-4,
                  models.lookup_m_eq_lut(re_, gd["T0"], Tmin, invdT, nT), 0.0, 1e-4])
    garr = _params_arr(gd)
    T4 = np.array([gd["T0"]] * 4)
    dt = 0.5e-15
    ts, fe, gdm, tev = [], [], [], []
    for n in range(1, 16001):
        tt = n * dt
        u = (tt - 0.4e-12) / 0.2e-12
        pabs = 4.0e21 * math.exp(-(u * u))
        m, out = _llb(oracle_lib, garr, lut, m, T4[0], T4[2], T4[3], dt)
        qtm = gd["Ms_TM"] * out[2] * (out[0] / dt)
        qre = gd["Ms_RE"] * out[3] * (out[1] / dt)
        oracle_lib.orc_update_4tm_cell(T4.ctypes.data_as(C.c_void_p), pabs, qtm, qre,
                                       garr.ctypes.data_as(C.c_void_p), dt)
        if n % 40 == 0:
            ts.append(tt * 1e12); fe.append(m[0]); gdm.append(m[3]); tev.append(T4[0])

    def zero_cross(t, y):
        for i in range(1, len(y)):
            if (y[i - 1] < 0) != (y[i] < 0):
                return t[i - 1] + (t[i] - t[i - 1]) * (0 - y[i - 1]) / (y[i] - y[i - 1])
        return float("nan")

    assert round(fe[0], 3) == 0.992 and round(fe[-1], 3) == -0.990
    assert round(gdm[0], 3) == -0.972 and round(gdm[-1], 3) == 0.968
    assert round(zero_cross(ts, fe), 2) == 0.92
    assert round(zero_cross(ts, gdm), 2) == 1.47
    assert round(max(tev)) == 2381


# ---- MO gyration (runtests.jl:581-603) --------------------------------------------------------
def _mo_patch(oracle_lib, qsign):
    from mphip import abi, dispersion, grid, models
    p = grid.FDTDParams()
    dt = 1.1e-17
    poles = dispersion.poles_array(dispersion.build_probe_mo_poles(dt, p.eps0, models.ferrimagnet_parameters()),
                                   np.float64)
    Ey = np.array([1.25])
    Ez = np.array([-0.42])
    arr = {k: np.zeros(2) for k in ("Pa", "Ja", "Pb", "Jb")}
    eo = np.zeros(1), np.zeros(1)
    idx = np.array([0], np.int32)
    fill = np.array([1.0])
    mpos = np.array([0], np.int32)
    inv_eps = np.array([1 / (p.eps0 * 4.0)])
    mtm, mre = np.array([qsign]), np.array([0.0])
    mo = abi.Mo()
    mo.n, mo.npoles, mo.dtype = 1, 2, abi.MP_F64
    mo.Pa_from_b, mo.Ja_from_b, mo.Pb_from_a, mo.Jb_from_a = (arr[k].ctypes.data for k in ("Pa", "Ja", "Pb", "Jb"))
    mo.E_old_a, mo.E_old_b = eo[0].ctypes.data, eo[1].ctypes.data
    mo.idx, mo.fill, mo.mpos, mo.poles = idx.ctypes.data, fill.ctypes.data, mpos.ctypes.data, poles.ctypes.data
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = oracle_lib.orc_patch_E_mo_gyration(vp(Ey), vp(Ez), C.byref(mo), vp(inv_eps), vp(inv_eps), abi.MP_F64,
                                            vp(mtm), vp(mre), 0.02, 0.0, dt)
    assert rc == 0
    return Ey[0], Ez[0]


def test_mo_gyration_even_in_m(oracle_lib):
    ey1, ez1 = _mo_patch(oracle_lib, 1.0)
    ey2, ez2 = _mo_patch(oracle_lib, -1.0)
    assert ey1 ** 2 + ez1 ** 2 == pytest.approx(ey2 ** 2 + ez2 ** 2, rel=1e-12, abs=1e-12)
    # the gyration must actually rotate the field (a dropped term would leave it unchanged)
    assert (ey1, ez1) != (1.25, -0.42) and (ey1, ez1) != (ey2, ez2)


# ---- ADE (runtests.jl:167-182 pins the recursion; here the kernel itself) ------------------
def _ade_drive(oracle_lib, pole, dt, omega, inv_eps, steps):
    """Drive one ADE cell with E = cos(w n dt) written before each patch; return P."""
    from mphip import abi
    poles = np.array(pole, np.float64)
    E = np.zeros(1)
    P, J, Eo = np.zeros(1), np.zeros(1), np.zeros(1)
    idx = np.array([0], np.int32)
    fill, inv = np.array([1.0]), np.array([inv_eps])
    a = abi.Ade()
    a.n, a.npoles, a.dtype, a.state_dtype, a.state_scale = 1, 1, abi.MP_F64, abi.MP_F64, 1.0
    a.P, a.J, a.E_old, a.idx, a.fill, a.inv_eps, a.poles = (x.ctypes.data for x in (P, J, Eo, idx, fill, inv, poles))
    Ps, Es = [], []
    for n in range(steps):
        E[0] = math.cos(omega * dt * n)
        rc = oracle_lib.orc_patch_E_dispersive(E.ctypes.data_as(C.c_void_p), C.byref(a), abi.MP_F64, dt)
        assert rc == 0
        Ps.append(P[0]); Es.append(E[0])
    return np.array(Ps), np.array(Es)


def test_ade_kernel_susceptibility_closed_form(oracle_lib):
    """With a vanishing screening factor the patch leaves E alone and P follows the
    pole: its steady-state response to cos(wt) is eps0*chi(w) with the continuous
    Lorentz chi = S / (w0^2 - w^2 - i g w), up to O((w dt)^2)."""
    from mphip import dispersion, grid
    eps0 = grid.EPS0
    w0, g, S = 4e15, 5e14, 2.0 * (4e15) ** 2
    dt = 2e-18
    omega = 2.0e15
    pole = dispersion.create_pole(w0, g, S, dt, eps0)
    n = 40000
    Ps, Es = _ade_drive(oracle_lib, pole, dt, omega, 1e-30, n)
    assert np.array_equal(Es, np.cos(omega * dt * np.arange(n)))
    k = np.arange(n // 2, n)
    t = k * dt
    M = np.column_stack([np.cos(omega * t), np.sin(omega * t), np.ones(len(k))])
    coef = np.linalg.lstsq(M, Ps[n // 2:], rcond=None)[0]
    chi = (coef[0] + 1j * coef[1]) / eps0
    chi_c = S / (w0 * w0 - omega * omega - 1j * g * omega)
    assert abs(chi - chi_c) / abs(chi_c) < 1e-3


def test_ade_kernel_constitutive_invariant(oracle_lib):
    """After the semi-implicit patch, E_final = E_temp - inv_eps * P_new holds (the
    screened field carries the new polarisation with the right sign)."""
    from mphip import abi, dispersion, grid
    eps0 = grid.EPS0
    dt = 1.9e-17
    poles = [dispersion.create_pole(0.0, 1e14, (1.5e16) ** 2, dt, eps0),
             dispersion.create_pole(4e15, 5e14, 2.0 * (4e15) ** 2, dt, eps0)]
    pa = np.array([c for p in poles for c in p], np.float64)
    rng = np.random.default_rng(3)
    n = 64
    E = rng.normal(size=n) * 1e7
    Etemp = E.copy()
    P = rng.normal(size=2 * n) * 1e-5
    J = rng.normal(size=2 * n) * 1e9
    Eo = rng.normal(size=n) * 1e7
    idx = np.arange(n, dtype=np.int32)
    fill = rng.uniform(0.2, 1.0, n)
    inv = 1.0 / (eps0 * rng.uniform(1.0, 4.0, n))
    a = abi.Ade()
    a.n, a.npoles, a.dtype, a.state_dtype, a.state_scale = n, 2, abi.MP_F64, abi.MP_F64, 1.0
    a.P, a.J, a.E_old, a.idx, a.fill, a.inv_eps, a.poles = (x.ctypes.data for x in (P, J, Eo, idx, fill, inv, pa))
    assert oracle_lib.orc_patch_E_dispersive(E.ctypes.data_as(C.c_void_p), C.byref(a), abi.MP_F64, dt) == 0
    Pnew = P[:n] + P[n:]
    np.testing.assert_allclose(E, Etemp - inv * Pnew, rtol=1e-9, atol=1e-9 * np.abs(Etemp).max())
    np.testing.assert_array_equal(Eo, E)


# ---- Yee + CPML end to end (runtests.jl:96-107, 249-272; docs/src/em_fdtd.md:78-88) ----------
def _em_spec(n, dx, npml, courant, pulse, src_index, field="f64", psi="f32"):
    from mphip import geometry, grid, sources
    from mphip.spec import build_state_spec
    g = grid.Grid3D(*(grid.uniform_axis(0.0, n * dx, dx) for _ in range(3)))
    geo = geometry.rasterize(geometry.Scene(), g)
    dt = grid.cfl_dt(g, courant=courant)
    src = sources.SoftSource(pulse, "Ez", src_index)
    return build_state_spec(g, geo, dt, compute_precision=field, field_precision=field, source=src,
                            n_pml=npml, psi_precision=psi), g


def _energy(st, g):
    from mphip.fields import field_energy
    return field_energy({k: st.field(k) for k in ("ex", "ey", "ez", "hx", "hy", "hz")}, g)


def test_3d_cpml_energy_ratio_documented():
    """docs/src/em_fdtd.md:78-88: 25^3 at 40 nm, PML(8), courant 0.35, Ez point source at
    the centre, 4000 steps -> energy final/peak = 8.9e-10."""
    import driver
    from mphip import grid, sources
    c0 = grid.C0
    pulse = sources.GaussianPulse(amplitude=1.0, tau=2e-15, t0=8e-15, omega=2 * math.pi * c0 / 800e-9)
    spec, g = _em_spec(25, 40e-9, 8, 0.35, pulse, (12, 12, 12))
    assert g.shape == (25, 25, 25)
    st = driver.OracleState(spec)
    peak = 0.0
    for _ in range(4000):
        st.step()
        peak = max(peak, _energy(st, g))
    ratio = _energy(st, g) / peak
    assert float(f"{ratio:.1e}") == 8.9e-10


def test_3d_cpml_residual_vs_pec():
    """runtests.jl:261-272: 10^3 at 100 nm, PML(3), courant 0.35, 1000 steps (the reference
    spreads its point source over the 8 cells around the centre; here one cell):
    residual energy ratio < 1e-4 and < 1% of a PEC box's."""
    import driver
    from mphip import grid, sources
    pulse = sources.GaussianPulse(amplitude=1.0, tau=2e-15, t0=8e-15, omega=2 * math.pi * grid.C0 / 800e-9)

    def residual(npml, pec):
        spec, g = _em_spec(10, 1e-7, npml, 0.35, pulse, (5, 5, 5))
        st = driver.OracleState(spec)
        peak = 0.0
        for _ in range(1000):
            st.step()
            if pec:   # apply_boundary!(::PEC) (src/fdtd/Boundary.jl:90-96)
                for k in ("ex", "ey", "ez", "dx", "dy", "dz"):
                    a = st.f[k].reshape(g.shape, order="F")
                    a[0, :, :] = 0; a[-1, :, :] = 0; a[:, 0, :] = 0; a[:, -1, :] = 0; a[:, :, 0] = 0; a[:, :, -1] = 0
            peak = max(peak, _energy(st, g))
        return _energy(st, g) / peak

    r_pml = residual(3, False)
    r_pec = residual(0, True)
    assert r_pml < 1e-4
    assert r_pml < 0.01 * r_pec


def test_cpml_absorbs_vs_wall():
    """runtests.jl:96-107: 20^3 at 50 nm, courant 0.4, 1200 steps: CPML(6) leaves less
    than half the energy of the reflecting box."""
    import driver
    from mphip import grid, sources
    pulse = sources.GaussianPulse(amplitude=1.0, tau=3e-15, t0=12e-15, omega=2 * math.pi * grid.C0 / 800e-9)

    def final_energy(npml):
        spec, g = _em_spec(20, 5e-8, npml, 0.4, pulse, (9, 9, 9))
        st = driver.OracleState(spec).run(1200)
        return _energy(st, g)

    e_cpml, e_wall = final_energy(6), final_energy(0)
    assert math.isfinite(e_cpml) and e_cpml < 0.5 * e_wall
