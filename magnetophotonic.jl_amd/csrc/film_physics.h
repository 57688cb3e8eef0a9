// Per-cell film physics shared by the plain integrator and the surrogate step, FP64:
// brillouin, lookup_m_eq_lut, hot_weight, damp_radial_growth, llb_step
// (src/physics/Magnetization.jl:7-263) and update_Te_energy / update_4tm
// (src/physics/Thermal.jl:24-40).  The expression order follows the reference (no
// contraction).
#pragma once
#include "common.h"

namespace mp {

// Uniform constants of the film update.  film_consts() derives them once on the host from
// mp_ferri_params with the same IEEE double operations the per-cell code would apply to
// them (no contraction on either side), so every value is identical to computing it in the
// kernel.  The kernels read them from device memory through scalar loads, so no uniform
// double occupies a vector register pair and no kernel carries the 97-field struct as
// arguments.
struct FilmConsts {
    // update_Te_energy / update_4tm (Thermal.jl:24-40)
    double gamma_e, half_gamma_e, Gel, Ges_TM, Ges_RE, inv_Cl, T0, inv_Cs_TM, inv_Cs_RE;
    // llb_step (Magnetization.jl:105-263)
    double T_Curie, inv_T_Curie, alpha0_TM, alpha0_RE, two_alpha0_TM, two_alpha0_RE;
    double J_TMRE, kb, inv_mu_TM, inv_mu_RE, two_K_over_Ms_TM, two_K_over_Ms_RE;
    double twelve_J_TMTM, twelve_J_RERE, gamma_gyro, Hsw0, Tsw, inv_dTsw, Ms_TM, Ms_RE;
    // hot_weight centres: 0.92, 0.95, 0.90 and 0.88 x T_Curie
    double Tw_scatter_TM, Tw_scatter_RE, Tw_cool_TM, Tw_cool_RE;
    // brillouin(S, x) of each sublattice (0 = TM, 1 = RE): a = (2S+1)/(2S), b = 1/(2S) and
    // the small-x series coefficients (a^2 - b^2)/3 and (a^4 - b^4)/45
    double bri_a[2], bri_b[2], bri_c1[2], bri_c3[2];
    // 2S+1 when it is an integer in [2, 64] (every physical spin), else 0
    int32_t bri_pow[2];
};

// Device copy of the constants for `p` (cached per device; re-derived and uploaded only
// when the parameters change, after the device has drained).  NULL on error.
const FilmConsts *film_consts(const mp_ferri_params *p);

// q^p for a uniform integer p >= 1, by squaring.
__device__ __forceinline__ double ipow(double q, int p) {
    double r = 1.0;
    while (true) {
        if (p & 1) r *= q;
        p >>= 1;
        if (!p) return r;
        q *= q;
    }
}

// brillouin(S, x) (Magnetization.jl:7-22) with a, b and the series coefficients of the
// sublattice taken from the constants (k = 0: TM, 1: RE).
//
// B = a coth(ax) - b coth(bx).  The reference forms each coth as (e+1)/(e-1) from its own
// exp(2ax) and exp(2bx).  Since a = (2S+1) b, both come from one exponential when 2S+1 is
// an integer: with q = exp(-2|bx|) and q_a = q^(2S+1) = exp(-2|ax|),
//   B = sign(x) [a (1+q_a)(1-q) - b (1+q)(1-q_a)] / ((1-q_a)(1-q)),
// one exponential and one division instead of two of each, with q <= 1 so nothing
// overflows.  Where the reference's exp(2ax) overflows (2|ax| > 709.78, |ax| <= 500) it
// yields inf/inf = NaN; that is kept.  The results agree with the reference's to rounding.
__device__ __forceinline__ double brillouin(const FilmConsts &c, int k, double x) {
    if (!__builtin_isfinite(x)) return 0.0;
    const double a = c.bri_a[k];
    const double b = c.bri_b[k];
    if (__builtin_fabs(x) < 1e-4) return c.bri_c1[k] * x - c.bri_c3[k] * (x * x * x);
    const double ax = a * x;
    const double bx = b * x;
    const int p = c.bri_pow[k];
    const double aax = __builtin_fabs(ax);
    if (p > 0 && aax <= 354.89135644669199) {
        const double q = exp(-2.0 * __builtin_fabs(bx));
        const double qa = ipow(q, p);
        const double num = a * ((1.0 + qa) * (1.0 - q)) - b * ((1.0 + q) * (1.0 - qa));
        const double den = (1.0 - qa) * (1.0 - q);
        return jl_clamp(__builtin_copysign(num / den, x), -1.0, 1.0);
    }
    if (p > 0 && aax <= 500.0 && ax > 0.0) return __builtin_nan("");
    const double coth_ax = __builtin_fabs(ax) > 500.0 ? __builtin_copysign(1.0, ax) : (exp(2.0 * ax) + 1.0) / (exp(2.0 * ax) - 1.0);
    const double coth_bx = __builtin_fabs(bx) > 500.0 ? __builtin_copysign(1.0, bx) : (exp(2.0 * bx) + 1.0) / (exp(2.0 * bx) - 1.0);
    return jl_clamp(a * coth_ax - b * coth_bx, -1.0, 1.0);
}

template <typename L>
__device__ __forceinline__ double lut_lookup(const L *lut, double T, double T_min, double inv_dT, int32_t n) {
    const double pos = (T - T_min) * inv_dT;
    if (!__builtin_isfinite(pos)) return 0.0;
    const double pos_c = jl_clamp(pos, 0.0, (double)(n - 1));
    const double base = floor(pos_c);
    int64_t i0 = (int64_t)base + 1;
    i0 = i0 < 1 ? 1 : (i0 > n ? n : i0);
    const int64_t i1 = i0 + 1 < (int64_t)n ? i0 + 1 : (int64_t)n;
    const double frac = jl_clamp(pos_c - base, 0.0, 1.0);
    return (1.0 - frac) * (double)lut[i0 - 1] + frac * (double)lut[i1 - 1];
}

__device__ __forceinline__ double hot_weight(double T, double Tsw, double inv_dTsw) {
    return 1.0 / (1.0 + exp(-(T - Tsw) * inv_dTsw));
}

__device__ __forceinline__ void damp_radial(double mx, double my, double mz, double &dx, double &dy,
                                            double &dz, double mag, double limit, double width) {
    const double radial_dot = mx * dx + my * dy + mz * dz;
    if (radial_dot > 0.0 && mag > limit - width) {
        const double scale = jl_clamp((limit - mag) / width, 0.0, 1.0);
        dx = dx * scale;
        dy = dy * scale;
        dz = dz * scale;
    }
}

// llb_step (Magnetization.jl:105-263); m[0..5] in/out; returns dm_TM_x, dm_RE_x, Hani_TM_x, Hani_RE_x.
template <typename L>
__device__ __forceinline__ void llb_step(double m[6], double te_current, double ts_TM_old,
                                         double ts_RE_old, const FilmConsts &gd,
                                         const L *tm_lut, const L *re_lut, double lut_T_min,
                                         double lut_inv_dT, int32_t lut_N, double dt,
                                         int brillouin_iters, double out[4]) {
    const double mx_TM = m[0], my_TM = m[1], mz_TM = m[2], mx_RE = m[3], my_RE = m[4], mz_RE = m[5];
    const double ts_avg = 0.5 * (ts_TM_old + ts_RE_old);
    const double ratio_TM = ts_TM_old * gd.inv_T_Curie;
    const double ratio_RE = ts_RE_old * gd.inv_T_Curie;
    double m_eq_TM = lut_lookup<L>(tm_lut, ts_TM_old, lut_T_min, lut_inv_dT, lut_N);
    double m_eq_RE = lut_lookup<L>(re_lut, ts_RE_old, lut_T_min, lut_inv_dT, lut_N);
    const double chi_par_TM = ratio_TM < 1.0 ? gd.alpha0_TM / (gd.T_Curie * jl_max(1.0 - ratio_TM, 0.01))
                                             : gd.alpha0_TM / (gd.T_Curie * jl_max(ratio_TM - 1.0, 0.01));
    const double chi_par_RE = ratio_RE < 1.0 ? gd.alpha0_RE / (gd.T_Curie * jl_max(1.0 - ratio_RE, 0.01))
                                             : gd.alpha0_RE / (gd.T_Curie * jl_max(ratio_RE - 1.0, 0.01));
    const double alpha_par_TM = ratio_TM < 1.0 ? gd.two_alpha0_TM * jl_max(ratio_TM, 0.01) / 3.0 : gd.alpha0_TM;
    const double alpha_perp_TM = ratio_TM < 1.0 ? gd.alpha0_TM * (1.0 - jl_max(ratio_TM, 0.01) / 3.0) : gd.alpha0_TM;
    const double alpha_par_RE = ratio_RE < 1.0 ? gd.two_alpha0_RE * jl_max(ratio_RE, 0.01) / 3.0 : gd.alpha0_RE;
    const double alpha_perp_RE = ratio_RE < 1.0 ? gd.alpha0_RE * (1.0 - jl_max(ratio_RE, 0.01) / 3.0) : gd.alpha0_RE;
    const double m2_TM = mx_TM * mx_TM + my_TM * my_TM + mz_TM * mz_TM;
    const double m2_RE = mx_RE * mx_RE + my_RE * my_RE + mz_RE * mz_RE;
    const double mag_TM = sqrt(m2_TM);
    const double mag_RE = sqrt(m2_RE);
    const double exchange_softening = jl_max(0.35, 300.0 / jl_max(te_current, 300.0));
    const double J_eff_TMRE = gd.J_TMRE * exchange_softening;
    double hot_window = 0.0;
    if (te_current > gd.T_Curie) {
        const double u = (te_current - gd.Tsw) * gd.inv_dTsw;
        hot_window = exp(-0.5 * (u * u));
    }
    const double Hsw = gd.Hsw0 * hot_window;
    const double twelve_J_eff = 12.0 * J_eff_TMRE;
    const double inv_kb_ts_TM = 1.0 / (gd.kb * jl_max(ts_TM_old, 1.0));
    const double inv_kb_ts_RE = 1.0 / (gd.kb * jl_max(ts_RE_old, 1.0));
    const double c_TM = twelve_J_eff * gd.inv_mu_TM;
    const double c_RE = twelve_J_eff * gd.inv_mu_RE;
    const double Hex_TM_x = c_TM * mx_RE, Hex_TM_y = c_TM * my_RE, Hex_TM_z = c_TM * mz_RE;
    const double Hex_RE_x = c_RE * mx_TM, Hex_RE_y = c_RE * my_TM, Hex_RE_z = c_RE * mz_TM;
    const double Hani_TM_x = gd.two_K_over_Ms_TM * mx_TM;
    const double Hani_RE_x = gd.two_K_over_Ms_RE * mx_RE;
    const double Hx_TM = Hex_TM_x + Hani_TM_x - Hsw, Hy_TM = Hex_TM_y, Hz_TM = Hex_TM_z;
    const double Hx_RE = Hex_RE_x + Hani_RE_x + Hsw, Hy_RE = Hex_RE_y, Hz_RE = Hex_RE_z;
    double Hlong_fac_TM, Hlong_fac_RE;
    if (ts_avg < gd.T_Curie) {
        for (int it = 0; it < brillouin_iters; ++it) {
            const double x_TM = (gd.twelve_J_TMTM * m_eq_TM + twelve_J_eff * m_eq_RE) * inv_kb_ts_TM;
            const double x_RE = (gd.twelve_J_RERE * m_eq_RE + twelve_J_eff * m_eq_TM) * inv_kb_ts_RE;
            m_eq_TM = jl_clamp(brillouin(gd, 0, x_TM), -1.0, 1.0);
            m_eq_RE = jl_clamp(brillouin(gd, 1, x_RE), -1.0, 1.0);
        }
        const double m_eq_floor = 0.01;
        Hlong_fac_TM = (__builtin_fabs(m_eq_TM) - mag_TM) / (2.0 * chi_par_TM * jl_max(__builtin_fabs(m_eq_TM), m_eq_floor));
        Hlong_fac_RE = (__builtin_fabs(m_eq_RE) - mag_RE) / (2.0 * chi_par_RE * jl_max(__builtin_fabs(m_eq_RE), m_eq_floor));
    } else {
        Hlong_fac_TM = -mag_TM / (2.0 * chi_par_TM);
        Hlong_fac_RE = -mag_RE / (2.0 * chi_par_RE);
    }
    const double mdotH_TM = mx_TM * Hx_TM + my_TM * Hy_TM + mz_TM * Hz_TM;
    const double mdotH_RE = mx_RE * Hx_RE + my_RE * Hy_RE + mz_RE * Hz_RE;
    const double cx_TM = my_TM * Hz_TM - mz_TM * Hy_TM;
    const double cy_TM = mz_TM * Hx_TM - mx_TM * Hz_TM;
    const double cz_TM = mx_TM * Hy_TM - my_TM * Hx_TM;
    const double cx_RE = my_RE * Hz_RE - mz_RE * Hy_RE;
    const double cy_RE = mz_RE * Hx_RE - mx_RE * Hz_RE;
    const double cz_RE = mx_RE * Hy_RE - my_RE * Hx_RE;
    const double ccx_TM = my_TM * cz_TM - mz_TM * cy_TM;
    const double ccy_TM = mz_TM * cx_TM - mx_TM * cz_TM;
    const double ccz_TM = mx_TM * cy_TM - my_TM * cx_TM;
    const double ccx_RE = my_RE * cz_RE - mz_RE * cy_RE;
    const double ccy_RE = mz_RE * cx_RE - mx_RE * cz_RE;
    const double ccz_RE = mx_RE * cy_RE - my_RE * cx_RE;
    const double inv_m2_TM = 1.0 / (m2_TM + 1e-30);
    const double inv_m2_RE = 1.0 / (m2_RE + 1e-30);
    const double g = gd.gamma_gyro;
    const double Lambda_TM = jl_clamp(g * Hlong_fac_TM, -1.0 / 0.12e-12, 1.0 / 2.80e-12);
    const double Lambda_RE = jl_clamp(g * Hlong_fac_RE, -1.0 / 0.45e-12, 1.0 / 3.20e-12);
    const double dm_TM_x_trans = -g * cx_TM + g * alpha_par_TM * inv_m2_TM * mdotH_TM * mx_TM - g * alpha_perp_TM * inv_m2_TM * ccx_TM;
    const double dm_TM_y_trans = -g * cy_TM + g * alpha_par_TM * inv_m2_TM * mdotH_TM * my_TM - g * alpha_perp_TM * inv_m2_TM * ccy_TM;
    const double dm_TM_z_trans = -g * cz_TM + g * alpha_par_TM * inv_m2_TM * mdotH_TM * mz_TM - g * alpha_perp_TM * inv_m2_TM * ccz_TM;
    const double dm_RE_x_trans = -g * cx_RE + g * alpha_par_RE * inv_m2_RE * mdotH_RE * mx_RE - g * alpha_perp_RE * inv_m2_RE * ccx_RE;
    const double dm_RE_y_trans = -g * cy_RE + g * alpha_par_RE * inv_m2_RE * mdotH_RE * my_RE - g * alpha_perp_RE * inv_m2_RE * ccy_RE;
    const double dm_RE_z_trans = -g * cz_RE + g * alpha_par_RE * inv_m2_RE * mdotH_RE * mz_RE - g * alpha_perp_RE * inv_m2_RE * ccz_RE;
    const double W_TM_scatter = hot_weight(te_current, gd.Tw_scatter_TM, 1.0 / 40.0);
    const double W_RE_scatter = hot_weight(ts_RE_old, gd.Tw_scatter_RE, 1.0 / 75.0);
    const double cool_TM = 1.0 - hot_weight(ts_TM_old, gd.Tw_cool_TM, 1.0 / 55.0);
    const double cool_RE = 1.0 - hot_weight(ts_RE_old, gd.Tw_cool_RE, 1.0 / 75.0);
    const double tm_switched_gate = hot_weight(-mx_TM, 0.006, 1.0 / 0.010);
    const double re_res_floor = 0.28 * W_TM_scatter * (1.0 - 0.50 * W_RE_scatter);
    const double tm_cross_floor = 0.16 * W_TM_scatter;
    const double tm_target_mag = 0.50 + 0.49 * cool_TM;
    const double tm_target_x_pre = -jl_max(jl_max(__builtin_fabs(mx_RE), re_res_floor), tm_cross_floor);
    const double tm_target_x_post = -tm_target_mag;
    const double tm_target_x = (1.0 - tm_switched_gate) * tm_target_x_pre + tm_switched_gate * tm_target_x_post;
    const double rate_TM_pre = 10.0e12 * W_TM_scatter;
    const double rate_TM_post = 1.15e12 * W_TM_scatter + 0.24e12 * cool_TM;
    const double rate_TM = (1.0 - tm_switched_gate) * rate_TM_pre + tm_switched_gate * rate_TM_post;
    const double re_branch_gate = tm_switched_gate * W_RE_scatter;
    const double re_target_mag = 0.18 + 0.80 * cool_RE;
    const double re_target_x = (1.0 - re_branch_gate) * mx_RE + re_branch_gate * re_target_mag;
    const double rate_RE = 0.45e12 * W_RE_scatter + 0.25e12 * cool_RE;
    double dm_TM_x = dm_TM_x_trans + rate_TM * (tm_target_x - mx_TM);
    double dm_TM_y = dm_TM_y_trans - rate_TM * my_TM;
    double dm_TM_z = dm_TM_z_trans - rate_TM * mz_TM;
    double dm_RE_x = dm_RE_x_trans + rate_RE * (re_target_x - mx_RE);
    double dm_RE_y = dm_RE_y_trans - rate_RE * my_RE;
    double dm_RE_z = dm_RE_z_trans - rate_RE * mz_RE;
    damp_radial(mx_TM, my_TM, mz_TM, dm_TM_x, dm_TM_y, dm_TM_z, mag_TM, 0.990, 0.55);
    damp_radial(mx_RE, my_RE, mz_RE, dm_RE_x, dm_RE_y, dm_RE_z, mag_RE, 0.968, 0.50);
    double lambda_TM_eff = Lambda_TM;
    double lambda_RE_eff = Lambda_RE;
    if (mx_TM < -0.020 && lambda_TM_eff < 0.0) lambda_TM_eff = 0.0;
    if (lambda_TM_eff > 0.0) lambda_TM_eff *= jl_clamp((0.990 - mag_TM) / 0.55, 0.0, 1.0);
    if (lambda_RE_eff > 0.0) lambda_RE_eff *= jl_clamp((0.968 - mag_RE) / 0.50, 0.0, 1.0);
    const double exp_TM = exp(jl_clamp(lambda_TM_eff * dt, -20.0, 20.0));
    const double exp_RE = exp(jl_clamp(lambda_RE_eff * dt, -20.0, 20.0));
    double m_TM_x_new = mx_TM * exp_TM + dt * dm_TM_x;
    double m_TM_y_new = my_TM * exp_TM + dt * dm_TM_y;
    double m_TM_z_new = mz_TM * exp_TM + dt * dm_TM_z;
    double m_RE_x_new = mx_RE * exp_RE + dt * dm_RE_x;
    double m_RE_y_new = my_RE * exp_RE + dt * dm_RE_y;
    double m_RE_z_new = mz_RE * exp_RE + dt * dm_RE_z;
    const double mag_new_TM = sqrt(m_TM_x_new * m_TM_x_new + m_TM_y_new * m_TM_y_new + m_TM_z_new * m_TM_z_new);
    if (mag_new_TM > 0.992) {
        const double s = 0.992 / mag_new_TM;
        m_TM_x_new *= s; m_TM_y_new *= s; m_TM_z_new *= s;
    }
    const double mag_new_RE = sqrt(m_RE_x_new * m_RE_x_new + m_RE_y_new * m_RE_y_new + m_RE_z_new * m_RE_z_new);
    if (mag_new_RE > 0.972) {
        const double s = 0.972 / mag_new_RE;
        m_RE_x_new *= s; m_RE_y_new *= s; m_RE_z_new *= s;
    }
    out[0] = m_TM_x_new - mx_TM;
    out[1] = m_RE_x_new - mx_RE;
    out[2] = Hani_TM_x;
    out[3] = Hani_RE_x;
    m[0] = m_TM_x_new; m[1] = m_TM_y_new; m[2] = m_TM_z_new;
    m[3] = m_RE_x_new; m[4] = m_RE_y_new; m[5] = m_RE_z_new;
}

// update_Te_energy + update_4tm (Thermal.jl:24-40)
__device__ __forceinline__ void update_4tm(double &te, double &tl, double &tsT, double &tsR, double pabs,
                                           double QT, double QR, const FilmConsts &gd, double dt) {
    const double tau_th = 2.0e-12;
    const double te0 = te, tl0 = tl, t0 = tsT, r0 = tsR;
    const double ue = gd.half_gamma_e * (te0 * te0);
    const double q = -gd.Gel * (te0 - tl0) - gd.Ges_TM * (te0 - t0) - gd.Ges_RE * (te0 - r0) + pabs;
    const double ue_new = jl_max(gd.half_gamma_e, ue + dt * q);
    te = sqrt(2.0 * ue_new / gd.gamma_e);
    tl = tl0 + dt * (gd.Gel * (te0 - tl0) * gd.inv_Cl) - dt * (tl0 - gd.T0) / tau_th;
    tsT = jl_max(1.0, t0 + dt * (gd.Ges_TM * (te0 - t0) - QT) * gd.inv_Cs_TM);
    tsR = jl_max(1.0, r0 + dt * (gd.Ges_RE * (te0 - r0) - QR) * gd.inv_Cs_RE);
}

}  // namespace mp
