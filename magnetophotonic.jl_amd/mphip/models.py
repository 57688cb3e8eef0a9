"""GdFeCo ferrimagnet parameters and the Brillouin equilibrium LUT (host side).

Restates src/physics/Models.jl:5-226 (`FerrimagnetParameters` with its 97 derived
fields, in declaration order = the C struct mp_ferri_params) and
src/physics/Magnetization.jl:7-86 (`brillouin`, `build_m_eq_lut`, `lookup_m_eq_lut`).
"""
from __future__ import annotations

import math

import numpy as np

# Field order of FerrimagnetParameters{T} (Models.jl:5-103) == mp_ferri_params.
FIELDS = (
    "gamma_e", "Cl", "Cs_TM", "Cs_RE", "Gel", "Ges_TM", "Ges_RE", "Gamma_E",
    "Hsw0", "Tsw", "dTsw", "gamma_gyro", "Ms_TM", "Ms_RE", "mu_TM", "mu_RE",
    "alpha0_TM", "alpha0_RE", "J_TMRE", "J_TMTM", "J_RERE", "K_TM", "K_RE",
    "T_Curie", "T_comp", "T0", "TcT", "TcR", "inv_TcT", "inv_TcR",
    "beta_T", "beta_R", "floor_T", "floor_R", "val300_T", "val300_R", "inv_val300_T", "inv_val300_R",
    "GammaT0", "GammaT_hot", "GammaR0", "GammaR_hot", "Tsw_long", "dTsw_long", "inv_dTsw_long",
    "ellT", "ellR", "inv_ellT", "inv_ellR", "GammaEx", "eta_quench", "inv_eta_quench",
    "GammaT_tail", "eta_tail", "inv_eta_tail", "mTM_tail_target", "field_to_K",
    "GTlT", "GTlR", "GsTR", "S_TM", "S_RE", "Q_voigt_TM", "Q_voigt_RE",
    "omega_pump", "eps_imag_pump", "eps_real_pump", "omega_probe", "eps_imag_probe", "eps_real_probe",
    "kb", "moment_ratio_TM_RE", "inv_Ms_sum",
    "inv_T_Curie", "inv_Cl", "inv_Cs_TM", "inv_Cs_RE",
    "ferri_RE_scale", "re_exchange_lag", "inv_mu_TM", "inv_mu_RE",
    "inv_Ms_TM", "inv_Ms_RE", "twelve_J_TMTM", "twelve_J_RERE",
    "two_K_over_Ms_TM", "two_K_over_Ms_RE",
    "inv_tau_TM_demag", "inv_tau_RE_demag", "moment_ratio_RE_TM",
    "HK_T_field", "HK_R_field", "HAF_T_field", "HAF_R_field", "alphaT_hot", "alphaR_hot", "seed_tilt",
)
assert len(FIELDS) == 97


class FerrimagnetParameters(dict):
    """FerrimagnetParameters{Float64} as an ordered name -> value map.

    `as_array(dtype)` packs the struct for the C-ABI; with dtype float32 each field is
    rounded to Float32 and widened back, which is what convert_compute_params(Float32, gd)
    hands the FP64 multiphysics arithmetic in the reference (Models.jl:246-252).
    """

    def as_array(self, compute_dtype=np.float64) -> np.ndarray:
        v = np.array([self[k] for k in FIELDS], dtype=np.float64)
        return v.astype(compute_dtype).astype(np.float64)

    def converted(self, compute_dtype) -> "FerrimagnetParameters":
        arr = self.as_array(compute_dtype)
        return FerrimagnetParameters(zip(FIELDS, (float(x) for x in arr)))


def ferrimagnet_parameters(
        c0=299792458.0, kb=1.380649e-23, muB=9.2740100783e-24, gamma_e=300.0, Cl=2.30e6,
        Gel=6.0e17, T0=300.0, T_Curie=540.0, T_comp=250.0, TcT=620.0, TcR=760.0, beta_T=0.50,
        beta_R=0.50, floor_T=0.05, floor_R=0.05, Ms_TM=1.10e6, Ms_RE=0.55e6, mu_TM=None,
        mu_RE=None, S_TM=1.0, S_RE=3.5, tau_TM_demag=100e-15, tau_RE_demag=430e-15, Cs_TM=5.0e4,
        Cs_RE=2.0e4, Ges_TM=None, Ges_RE=None, Gamma_E=1.0 / 140e-15, gamma_gyro=1.76085963023e11,
        alpha0_TM=0.05, alpha0_RE=0.05, J_TMRE=-4.77e-21, J_TMTM=3.46e-21, J_RERE=1.389e-21,
        K_total=3.8e4, K_TM=None, K_RE=None, Hsw0=0.0, Tsw=None, dTsw=50.0, Tsw_long=None,
        dTsw_long=55.0, GammaT0=1.0 / 5.0e-12, GammaT_hot=None, GammaR0=1.0 / 20.0e-12,
        GammaR_hot=None, GammaEx=None, eta_quench=1.0, GammaT_tail=0.0, eta_tail=1.0,
        mTM_tail_target=0.0, GTlT=None, GTlR=None, GsTR=0.90e17, lambda_pump=800e-9,
        lambda_probe=532e-9, n_pump=3.50, k_pump=3.73, n_probe=2.73, k_probe=4.60,
        Q_voigt_TM=0.020, Q_voigt_RE=0.006, ellT=1.0, ellR=None, field_to_K=None,
        ferri_RE_scale=None, re_exchange_lag=None, seed_tilt=1e-4) -> FerrimagnetParameters:
    """FerrimagnetParameters{T}(; kwargs...) defaults (Models.jl:107-226), GdFeCo preset."""
    if mu_TM is None:
        mu_TM = 1.92 * muB
    if mu_RE is None:
        mu_RE = 7.63 * muB
    Ges_TM_v = Cs_TM / tau_TM_demag if Ges_TM is None else float(Ges_TM)
    Ges_RE_v = Cs_RE / tau_RE_demag if Ges_RE is None else float(Ges_RE)
    K_TM_v = 0.5 * K_total if K_TM is None else float(K_TM)
    K_RE_v = 0.5 * K_total if K_RE is None else float(K_RE)
    Tsw_v = float(T_Curie) if Tsw is None else float(Tsw)
    Tsw_long_v = float(T_Curie) if Tsw_long is None else float(Tsw_long)
    GammaT_hot_v = 1.0 / tau_TM_demag if GammaT_hot is None else float(GammaT_hot)
    GammaR_hot_v = 1.0 / tau_RE_demag if GammaR_hot is None else float(GammaR_hot)
    GammaEx_v = float(Gamma_E) if GammaEx is None else float(GammaEx)
    ellR_v = mu_RE / mu_TM if ellR is None else float(ellR)
    field_to_K_v = mu_TM / kb if field_to_K is None else float(field_to_K)
    GTlT_v = Ges_TM_v if GTlT is None else float(GTlT)
    GTlR_v = Ges_RE_v if GTlR is None else float(GTlR)
    ferri_RE_scale_v = Ms_RE / Ms_TM if ferri_RE_scale is None else float(ferri_RE_scale)
    re_exchange_lag_v = mu_TM / mu_RE if re_exchange_lag is None else float(re_exchange_lag)
    val300_T = max(float(floor_T), math.sqrt(max(0.0, 1.0 - T0 / TcT)))
    val300_R = max(float(floor_R), math.sqrt(max(0.0, 1.0 - T0 / TcR)))
    omega_pump = 2 * math.pi * c0 / lambda_pump
    omega_probe = 2 * math.pi * c0 / lambda_probe
    eps_real_pump = n_pump * n_pump - k_pump * k_pump
    eps_imag_pump = 2.0 * n_pump * k_pump
    eps_real_probe = n_probe * n_probe - k_probe * k_probe
    eps_imag_probe = 2.0 * n_probe * k_probe
    HK_T = 2.0 * K_TM_v / Ms_TM
    HK_R = 2.0 * K_RE_v / Ms_RE
    HAF_T = abs(J_TMRE) / mu_TM
    HAF_R = abs(J_TMRE) / mu_RE
    vals = (
        gamma_e, Cl, Cs_TM, Cs_RE, Gel, Ges_TM_v, Ges_RE_v, Gamma_E,
        Hsw0, Tsw_v, dTsw, gamma_gyro, Ms_TM, Ms_RE, mu_TM, mu_RE,
        alpha0_TM, alpha0_RE, J_TMRE, J_TMTM, J_RERE, K_TM_v, K_RE_v,
        T_Curie, T_comp, T0, TcT, TcR, 1 / TcT, 1 / TcR,
        beta_T, beta_R, floor_T, floor_R, val300_T, val300_R, 1 / val300_T, 1 / val300_R,
        GammaT0, GammaT_hot_v, GammaR0, GammaR_hot_v, Tsw_long_v, dTsw_long, 1 / dTsw_long,
        ellT, ellR_v, 1 / ellT, 1 / ellR_v, GammaEx_v, eta_quench, 1 / eta_quench,
        GammaT_tail, eta_tail, 1 / eta_tail, mTM_tail_target, field_to_K_v,
        GTlT_v, GTlR_v, GsTR, S_TM, S_RE, Q_voigt_TM, Q_voigt_RE,
        omega_pump, eps_imag_pump, eps_real_pump, omega_probe, eps_imag_probe, eps_real_probe,
        kb, mu_TM / mu_RE, 1 / (Ms_TM + Ms_RE),
        1 / T_Curie, 1 / Cl, 1 / Cs_TM, 1 / Cs_RE,
        ferri_RE_scale_v, re_exchange_lag_v, 1 / mu_TM, 1 / mu_RE,
        1 / Ms_TM, 1 / Ms_RE, 12 * J_TMTM, 12 * J_RERE,
        2 * K_TM_v / Ms_TM, 2 * K_RE_v / Ms_RE,
        1 / tau_TM_demag, 1 / tau_RE_demag, mu_RE / mu_TM,
        HK_T, HK_R, HAF_T, HAF_R, alpha0_TM, alpha0_RE, seed_tilt,
    )
    return FerrimagnetParameters(zip(FIELDS, (float(v) for v in vals)))


def optical_coupling(gd, m_TM_x: float, m_RE_x: float) -> float:
    """optical_coupling (Models.jl:255-257): Voigt-weighted net moment."""
    return gd["Q_voigt_TM"] * float(m_TM_x) + gd["Q_voigt_RE"] * float(m_RE_x)


def _clamp(x, lo, hi):
    return hi if x > hi else (lo if x < lo else x)


def _div(a: float, b: float) -> float:
    """IEEE division (inf/inf = nan, x/0 = +-inf) without Python's ZeroDivisionError."""
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def _exp(x: float) -> float:
    """IEEE exp: +Inf on overflow (math.exp raises instead)."""
    try:
        return math.exp(x)
    except OverflowError:
        return math.inf


def brillouin(S: float, x: float) -> float:
    """brillouin (Magnetization.jl:7-22)."""
    S = float(S)
    x = float(x)
    if not math.isfinite(x):
        return 0.0
    a = (2.0 * S + 1.0) / (2.0 * S)
    b = 1.0 / (2.0 * S)
    if abs(x) < 1e-4:
        a2, b2 = a * a, b * b
        return ((a * a - b * b) / 3.0) * x - ((a2 * a2 - b2 * b2) / 45.0) * (x * x * x)
    ax = a * x
    bx = b * x
    coth_ax = math.copysign(1.0, ax) if abs(ax) > 500.0 else _div(_exp(2.0 * ax) + 1.0, _exp(2.0 * ax) - 1.0)
    coth_bx = math.copysign(1.0, bx) if abs(bx) > 500.0 else _div(_exp(2.0 * bx) + 1.0, _exp(2.0 * bx) - 1.0)
    return _clamp(a * coth_ax - b * coth_bx, -1.0, 1.0)


def build_m_eq_lut(gd, T_min: float = 1.0, T_max: float = 1800.0, dT: float = 1.0):
    """build_m_eq_lut (Magnetization.jl:50-73): mean-field m_eq(T) of both sublattices,
    16 fixed-point refinements per temperature.  Returns (tm, re, T_min, 1/dT, nT)."""
    nT = int(math.floor((T_max - T_min) / dT)) + 1
    tm = np.zeros(nT)
    re = np.zeros(nT)
    kb = gd["kb"]
    for i in range(nT):
        T = T_min + i * dT
        if T < gd["T_Curie"]:
            m0 = max(1.0 - T / gd["T_Curie"], 0.0) ** 0.365
            m_tm, m_re = m0, -m0
            J_eff = gd["J_TMRE"] * (300.0 / max(T, 300.0))
            for _ in range(16):
                x_TM = (12.0 * gd["J_TMTM"] * m_tm + 12.0 * J_eff * m_re) / (kb * max(T, 1.0))
                x_RE = (12.0 * gd["J_RERE"] * m_re + 12.0 * J_eff * m_tm) / (kb * max(T, 1.0))
                m_tm = _clamp(brillouin(gd["S_TM"], x_TM), -1.0, 1.0)
                m_re = _clamp(brillouin(gd["S_RE"], x_RE), -1.0, 1.0)
            tm[i] = m_tm
            re[i] = m_re
    return tm, re, T_min, 1.0 / dT, nT


def lookup_m_eq_lut(lut, T, T_min, inv_dT, lut_N) -> float:
    """lookup_m_eq_lut (Magnetization.jl:75-86): clamped linear interpolation."""
    pos = (float(T) - T_min) * inv_dT
    if not math.isfinite(pos):
        return 0.0
    pos_c = _clamp(pos, 0.0, float(lut_N - 1))
    base = math.floor(pos_c)
    i0 = _clamp(int(base) + 1, 1, int(lut_N))
    i1 = min(i0 + 1, int(lut_N))
    frac = _clamp(pos_c - base, 0.0, 1.0)
    return (1.0 - frac) * float(lut[i0 - 1]) + frac * float(lut[i1 - 1])
