/*
 * ORACLE (test infrastructure only).  CPU restatement of the surrogate's film side:
 * the training rows (features + rates of the plain integrator) and the surrogate step
 * given the network's rates.  The network itself is restated in oracle/mlp.py.
 * The plain substep is orc_llb_step + orc_update_4tm, the same code the plain
 * multiphysics restatement runs (src/fdtd/Kernels.jl:362-412).
 */

static void orc_cell_features(float f[11], double te, double tl, double tsT, double tsR, double mT,
                              double mR, double pabs, double dt, const mp_ferri_params *gd) {
    const double pn = pabs * 1e-18;
    const double apn = fabs(pn);
    const double ash = copysign(log(apn + sqrt(apn * apn + 1.0)), pn);
    f[0] = (float)te;
    f[1] = (float)tl;
    f[2] = (float)tsT;
    f[3] = (float)tsR;
    f[4] = (float)mT;
    f[5] = (float)mR;
    f[6] = (float)ash;
    f[7] = (float)log(dt * 1e18);
    f[8] = (float)gd->gamma_e;
    f[9] = (float)log(gd->Gel);
    f[10] = (float)gd->T_Curie;
}

#define ORC_SURR(LT)                                                                          \
static void orc_plain_substep_##LT(double *te, double *tl, double *tsT, double *tsR, double m[6], \
                                   double pabs, const mp_ferri_params *gd, const LT *tm_lut,   \
                                   const LT *re_lut, double T_min, double inv_dT, int32_t lut_N, \
                                   double sub_dt, int iters) {                                 \
    double out[4];                                                                            \
    orc_llb_step_##LT(m, *te, *tsT, *tsR, gd, tm_lut, re_lut, T_min, inv_dT, lut_N, sub_dt, iters, out); \
    double QT = gd->Ms_TM * out[2] * (out[0] / sub_dt);                                       \
    double QR = gd->Ms_RE * out[3] * (out[1] / sub_dt);                                       \
    orc_update_4tm(te, tl, tsT, tsR, pabs, QT, QR, gd, sub_dt);                               \
}                                                                                             \
static void orc_dataset_##LT(const mp_film *f, const mp_ferri_params *gd, const LT *tm_lut,   \
                             const LT *re_lut, double T_min, double inv_dT, int32_t lut_N,    \
                             double dt, int subcycles, int iters, int has_pabs, int64_t stride, \
                             int64_t rows, float *X, float *T) {                              \
    const LT *Te = (const LT *)f->Te, *Tl = (const LT *)f->Tl, *TsT = (const LT *)f->Ts_TM;   \
    const LT *TsR = (const LT *)f->Ts_RE, *mTx = (const LT *)f->m_TM_x;                       \
    const LT *mTy = (const LT *)f->m_TM_y, *mTz = (const LT *)f->m_TM_z;                      \
    const LT *mRx = (const LT *)f->m_RE_x, *mRy = (const LT *)f->m_RE_y, *mRz = (const LT *)f->m_RE_z; \
    _Pragma("omp parallel for schedule(static)")                                             \
    for (int64_t r = 0; r < rows; ++r) {                                                      \
        const int64_t n = r * stride;                                                         \
        const double pabs = has_pabs ? f->window_sum[n] / dt : 0.0;                           \
        double te = (double)Te[n], tl = (double)Tl[n], tsT = (double)TsT[n], tsR = (double)TsR[n]; \
        double m[6] = {(double)mTx[n], (double)mTy[n], (double)mTz[n],                        \
                       (double)mRx[n], (double)mRy[n], (double)mRz[n]};                       \
        orc_cell_features(X + r * 11, te, tl, tsT, tsR, m[0], m[3], pabs, dt, gd);            \
        const double te0 = te, tl0 = tl, tsT0 = tsT, tsR0 = tsR, mT0 = m[0], mR0 = m[3];      \
        const double sub_dt = dt / (double)subcycles;                                         \
        for (int s = 0; s < subcycles; ++s)                                                   \
            orc_plain_substep_##LT(&te, &tl, &tsT, &tsR, m, pabs, gd, tm_lut, re_lut, T_min,  \
                                   inv_dT, lut_N, sub_dt, iters);                             \
        const double hg = 0.5 * gd->gamma_e;                                                  \
        float *t = T + r * 6;                                                                 \
        t[0] = (float)((hg * (te * te) - hg * (te0 * te0)) / dt - pabs);                      \
        t[1] = (float)((tl - tl0) / dt);                                                      \
        t[2] = (float)((tsT - tsT0) / dt);                                                    \
        t[3] = (float)((tsR - tsR0) / dt);                                                    \
        t[4] = (float)((m[0] - mT0) / dt);                                                    \
        t[5] = (float)((m[3] - mR0) / dt);                                                    \
    }                                                                                         \
}                                                                                             \
static int64_t orc_surrogate_##LT(mp_film *f, const mp_ferri_params *gd, const LT *tm_lut,    \
                                  const LT *re_lut, double T_min, double inv_dT, int32_t lut_N, \
                                  const float *rates, double dt, int subcycles, int iters,    \
                                  int has_pabs, double tol) {                                 \
    LT *Te = (LT *)f->Te, *Tl = (LT *)f->Tl, *TsT = (LT *)f->Ts_TM, *TsR = (LT *)f->Ts_RE;    \
    LT *mTx = (LT *)f->m_TM_x, *mTy = (LT *)f->m_TM_y, *mTz = (LT *)f->m_TM_z;                \
    LT *mRx = (LT *)f->m_RE_x, *mRy = (LT *)f->m_RE_y, *mRz = (LT *)f->m_RE_z;                \
    int64_t fallbacks = 0;                                                                    \
    _Pragma("omp parallel for schedule(static) reduction(+:fallbacks)")                      \
    for (int64_t n = 0; n < f->n; ++n) {                                                      \
        double pabs = 0.0;                                                                    \
        if (has_pabs) {                                                                       \
            pabs = f->window_sum[n] / dt;                                                     \
            f->window_sum[n] = 0.0;                                                           \
        }                                                                                     \
        double te = (double)Te[n], tl = (double)Tl[n], tsT = (double)TsT[n], tsR = (double)TsR[n]; \
        double m[6] = {(double)mTx[n], (double)mTy[n], (double)mTz[n],                        \
                       (double)mRx[n], (double)mRy[n], (double)mRz[n]};                       \
        const float *r = rates + n * 6;                                                       \
        const double sub_dt = dt / (double)subcycles;                                         \
        /* midpoint check of the network's step against one plain substep */                  \
        int accept = tol >= 0.0;                                                              \
        const double hg = 0.5 * gd->gamma_e;                                                  \
        const double ue0 = hg * (te * te);                                                    \
        const double S = (double)subcycles;                                                   \
        double e_te = 0, e_tl = 0, e_tsT = 0, e_tsR = 0, e_m[6] = {0, 0, 0, 0, 0, 0};         \
        if (accept) {                                                                         \
            e_te = sqrt(jl_max(hg, ue0 + dt * (pabs + (double)r[0])) / hg);                   \
            e_tl = jl_max(1.0, tl + dt * (double)r[1]);                                       \
            e_tsT = jl_max(1.0, tsT + dt * (double)r[2]);                                     \
            e_tsR = jl_max(1.0, tsR + dt * (double)r[3]);                                     \
            e_m[0] = m[0] + dt * (double)r[4];                                                \
            e_m[3] = m[3] + dt * (double)r[5];                                                \
            accept = isfinite(e_te) && isfinite(e_tl) && isfinite(e_tsT) && isfinite(e_tsR) &&\
                     fabs(e_m[0]) <= 1.0 && fabs(e_m[3]) <= 1.0;                              \
        }                                                                                     \
        if (accept) {                                                                         \
            const double h_te = 0.5 * (te + e_te), h_tl = 0.5 * (tl + e_tl);                  \
            const double h_tsT = 0.5 * (tsT + e_tsT), h_tsR = 0.5 * (tsR + e_tsR);            \
            const double h_m[6] = {0.5 * (m[0] + e_m[0]), m[1], m[2], 0.5 * (m[3] + e_m[3]), m[4], m[5]};\
            double b_te = h_te, b_tl = h_tl, b_tsT = h_tsT, b_tsR = h_tsR;                    \
            double b_m[6];                                                                    \
            for (int c = 0; c < 6; ++c) b_m[c] = h_m[c];                                      \
            orc_plain_substep_##LT(&b_te, &b_tl, &b_tsT, &b_tsR, b_m, pabs, gd, tm_lut, re_lut, T_min,\
                                   inv_dT, lut_N, sub_dt, iters);                             \
            double res = fabs(e_te - te - S * (b_te - h_te)) / e_te;                          \
            res = jl_max(res, fabs(e_tl - tl - S * (b_tl - h_tl)) / e_tl);                    \
            res = jl_max(res, fabs(e_tsT - tsT - S * (b_tsT - h_tsT)) / e_tsT);               \
            res = jl_max(res, fabs(e_tsR - tsR - S * (b_tsR - h_tsR)) / e_tsR);               \
            res = jl_max(res, fabs(e_m[0] - m[0] - S * (b_m[0] - h_m[0])));                   \
            res = jl_max(res, fabs(e_m[3] - m[3] - S * (b_m[3] - h_m[3])));                   \
            e_m[1] = m[1] + S * (b_m[1] - h_m[1]);                                            \
            e_m[2] = m[2] + S * (b_m[2] - h_m[2]);                                            \
            e_m[4] = m[4] + S * (b_m[4] - h_m[4]);                                            \
            e_m[5] = m[5] + S * (b_m[5] - h_m[5]);                                            \
            accept = res <= tol;                                                              \
        }                                                                                     \
        if (accept) {                                                                         \
            te = e_te; tl = e_tl; tsT = e_tsT; tsR = e_tsR;                                   \
            for (int c = 0; c < 6; ++c) m[c] = e_m[c];                                        \
        } else {                                                                              \
            for (int k = 0; k < subcycles; ++k)                                               \
                orc_plain_substep_##LT(&te, &tl, &tsT, &tsR, m, pabs, gd, tm_lut, re_lut, T_min,\
                                       inv_dT, lut_N, sub_dt, iters);                         \
            fallbacks += 1;                                                                   \
        }                                                                                     \
        Te[n] = (LT)te; Tl[n] = (LT)tl; TsT[n] = (LT)tsT; TsR[n] = (LT)tsR;                   \
        mTx[n] = (LT)m[0]; mTy[n] = (LT)m[1]; mTz[n] = (LT)m[2];                              \
        mRx[n] = (LT)m[3]; mRy[n] = (LT)m[4]; mRz[n] = (LT)m[5];                              \
    }                                                                                         \
    return fallbacks;                                                                         \
}
ORC_SURR(float)
ORC_SURR(double)
