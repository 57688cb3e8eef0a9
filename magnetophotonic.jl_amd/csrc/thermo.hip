// 4-temperature bath + two-sublattice LLB per film cell, FP64 arithmetic on gfx950.
//
// Reference: multiphysics_kernel! (src/fdtd/Kernels.jl:362-412) and its launcher
// (:648-666); update_Te_energy / update_4tm (src/physics/Thermal.jl:24-40);
// brillouin, lookup_m_eq_lut, hot_weight, damp_radial_growth, llb_step
// (src/physics/Magnetization.jl:7-263).  One thread per film cell, `subcycles`
// substeps held in registers; state is read once and written once per call.  The
// expression order follows the reference (no contraction).
#include "film_physics.h"

namespace mp {

template <typename L, bool HAS_PABS>
__global__ __launch_bounds__(128) void multiphysics_kernel(
    L *__restrict__ Te, L *__restrict__ Tl, L *__restrict__ TsT, L *__restrict__ TsR,
    L *__restrict__ mTx, L *__restrict__ mTy, L *__restrict__ mTz,
    L *__restrict__ mRx, L *__restrict__ mRy, L *__restrict__ mRz,
    double *__restrict__ window, const mp_ferri_params gd, const L *__restrict__ tm_lut,
    const L *__restrict__ re_lut, double T_min, double inv_dT, int32_t lut_N, double dt,
    int subcycles, int brillouin_iters, int64_t N) {
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += (int64_t)gridDim.x * blockDim.x) {
    double pabs = 0.0;
    if (HAS_PABS) {
        pabs = window[n] / dt;
        window[n] = 0.0;
    }
    double te = (double)Te[n], tl = (double)Tl[n], tsT = (double)TsT[n], tsR = (double)TsR[n];
    double m[6] = {(double)mTx[n], (double)mTy[n], (double)mTz[n], (double)mRx[n], (double)mRy[n], (double)mRz[n]};
    const double sub_dt = dt / (double)subcycles;
    for (int s = 0; s < subcycles; ++s) {
        double out[4];
        llb_step<L>(m, te, tsT, tsR, gd, tm_lut, re_lut, T_min, inv_dT, lut_N, sub_dt, brillouin_iters, out);
        const double Q_spin_TM = gd.Ms_TM * out[2] * (out[0] / sub_dt);
        const double Q_spin_RE = gd.Ms_RE * out[3] * (out[1] / sub_dt);
        update_4tm(te, tl, tsT, tsR, pabs, Q_spin_TM, Q_spin_RE, gd, sub_dt);
    }
    Te[n] = (L)te; Tl[n] = (L)tl; TsT[n] = (L)tsT; TsR[n] = (L)tsR;
    mTx[n] = (L)m[0]; mTy[n] = (L)m[1]; mTz[n] = (L)m[2];
    mRx[n] = (L)m[3]; mRy[n] = (L)m[4]; mRz[n] = (L)m[5];
  }
}

// Workgroups of the film update: at most 1024 (MP_FILM_GRID overrides; 0 = one per 128
// cells).  The kernel loops over the cells; on the 2.6 M-cell film a 1024-workgroup grid
// overlaps the Yee updates of the next step with 1% less step time than a full grid.
static int64_t film_grid(int64_t N) {
    static int64_t cap = -1;
    if (cap < 0) {
        const char *e = getenv("MP_FILM_GRID");
        cap = e ? atoll(e) : 1024;
    }
    const int64_t full = cdiv(N, 128);
    return cap > 0 && cap < full ? cap : full;
}

template <typename L>
static int launch_mp(mp_film *f, const mp_ferri_params *p, const mp_lut *lut, double dt,
                     int subcycles, int iters, int has_pabs) {
    const int64_t N = f->n;
    // The launcher rounds dt, T_min and inv_dT to the compute type (Kernels.jl:662-663).
    const double dtc = (double)(L)dt, tmin = (double)(L)lut->T_min, invdt = (double)(L)lut->inv_dT;
    KernelTimer t("multiphysics");
#define MP_ARGS (L *)f->Te, (L *)f->Tl, (L *)f->Ts_TM, (L *)f->Ts_RE, (L *)f->m_TM_x, (L *)f->m_TM_y, \
                (L *)f->m_TM_z, (L *)f->m_RE_x, (L *)f->m_RE_y, (L *)f->m_RE_z, f->window_sum, *p,   \
                (const L *)lut->tm, (const L *)lut->re, tmin, invdt, lut->n, dtc, subcycles, iters, N
    if (has_pabs) multiphysics_kernel<L, true><<<film_grid(N), 128, 0, stream()>>>(MP_ARGS);
    else multiphysics_kernel<L, false><<<film_grid(N), 128, 0, stream()>>>(MP_ARGS);
#undef MP_ARGS
    MP_LAUNCH_CHECK("multiphysics_kernel");
    return MP_OK;
}

}  // namespace mp

using namespace mp;

MP_API int mp_multiphysics_step(mp_film *f, const mp_ferri_params *p, const mp_lut *lut, double dt,
                                int32_t subcycles, int32_t brillouin_iters, int32_t has_pabs) {
    if (!f || !p || !lut) return fail_arg("mp_multiphysics_step: null argument");
    if (f->n <= 0) return MP_OK;  // Kernels.jl:654
    if (subcycles <= 0) return fail_arg("subcycles must be positive");  // Coupling.jl:30
    if (lut->n < 1 || !lut->tm || !lut->re) return fail_arg("mp_multiphysics_step: empty LUT");
    if (lut->dtype != f->dtype) return fail_arg("mp_multiphysics_step: LUT dtype must equal film dtype");
    if (has_pabs && !f->window_sum) return fail_arg("mp_multiphysics_step: missing window_sum");
    if (f->dtype == MP_F32) return launch_mp<float>(f, p, lut, dt, subcycles, brillouin_iters, has_pabs);
    if (f->dtype == MP_F64) return launch_mp<double>(f, p, lut, dt, subcycles, brillouin_iters, has_pabs);
    return fail_arg("mp_multiphysics_step: unsupported film dtype %d", f->dtype);
}
