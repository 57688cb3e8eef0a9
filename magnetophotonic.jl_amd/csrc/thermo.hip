// 4-temperature bath + two-sublattice LLB per film cell, FP64 arithmetic on gfx950.
//
// Reference: multiphysics_kernel! (src/fdtd/Kernels.jl:362-412) and its launcher
// (:648-666); update_Te_energy / update_4tm (src/physics/Thermal.jl:24-40);
// brillouin, lookup_m_eq_lut, hot_weight, damp_radial_growth, llb_step
// (src/physics/Magnetization.jl:7-263).  One thread per film cell, `subcycles`
// substeps held in registers; state is read once and written once per call.  The
// expression order follows the reference (no contraction).
#include <cmath>
#include <cstring>
#include <mutex>

#include "film_physics.h"

namespace mp {

static FilmConsts derive_consts(const mp_ferri_params &p) {
    FilmConsts c;
    c.gamma_e = p.gamma_e;
    c.half_gamma_e = 0.5 * p.gamma_e;
    c.Gel = p.Gel;
    c.Ges_TM = p.Ges_TM;
    c.Ges_RE = p.Ges_RE;
    c.inv_Cl = p.inv_Cl;
    c.T0 = p.T0;
    c.inv_Cs_TM = p.inv_Cs_TM;
    c.inv_Cs_RE = p.inv_Cs_RE;
    c.T_Curie = p.T_Curie;
    c.inv_T_Curie = p.inv_T_Curie;
    c.alpha0_TM = p.alpha0_TM;
    c.alpha0_RE = p.alpha0_RE;
    c.two_alpha0_TM = p.alpha0_TM * 2.0;
    c.two_alpha0_RE = p.alpha0_RE * 2.0;
    c.J_TMRE = p.J_TMRE;
    c.kb = p.kb;
    c.inv_mu_TM = p.inv_mu_TM;
    c.inv_mu_RE = p.inv_mu_RE;
    c.two_K_over_Ms_TM = p.two_K_over_Ms_TM;
    c.two_K_over_Ms_RE = p.two_K_over_Ms_RE;
    c.twelve_J_TMTM = p.twelve_J_TMTM;
    c.twelve_J_RERE = p.twelve_J_RERE;
    c.gamma_gyro = p.gamma_gyro;
    c.Hsw0 = p.Hsw0;
    c.Tsw = p.Tsw;
    c.inv_dTsw = 1.0 / p.dTsw;
    c.Ms_TM = p.Ms_TM;
    c.Ms_RE = p.Ms_RE;
    c.Tw_scatter_TM = 0.92 * p.T_Curie;
    c.Tw_scatter_RE = 0.95 * p.T_Curie;
    c.Tw_cool_TM = 0.90 * p.T_Curie;
    c.Tw_cool_RE = 0.88 * p.T_Curie;
    const double S[2] = {p.S_TM, p.S_RE};
    for (int k = 0; k < 2; ++k) {
        const double a = (2.0 * S[k] + 1.0) / (2.0 * S[k]);
        const double b = 1.0 / (2.0 * S[k]);
        const double a2 = a * a, b2 = b * b;
        c.bri_a[k] = a;
        c.bri_b[k] = b;
        c.bri_c1[k] = (a * a - b * b) / 3.0;
        c.bri_c3[k] = (a2 * a2 - b2 * b2) / 45.0;
        const double p = 2.0 * S[k] + 1.0;
        c.bri_pow[k] = (p == std::floor(p) && p >= 2.0 && p <= 64.0) ? (int32_t)p : 0;
    }
    return c;
}

namespace {
struct ConstsCache {
    bool valid = false;
    mp_ferri_params last;
    FilmConsts *dev = nullptr;
};
constexpr int kMaxDevices = 64;
ConstsCache g_consts[kMaxDevices];
std::mutex g_consts_mu;
}  // namespace

const FilmConsts *film_consts(const mp_ferri_params *p) {
    int dev = 0;
    if (check_hip(hipGetDevice(&dev), "hipGetDevice")) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        fail_arg("film_consts: device %d out of range", dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_consts_mu);
    ConstsCache &cc = g_consts[dev];
    if (cc.valid && std::memcmp(&cc.last, p, sizeof(*p)) == 0) return cc.dev;
    const FilmConsts h = derive_consts(*p);
    if (!cc.dev) {
        if (check_hip(hipMalloc((void **)&cc.dev, sizeof(FilmConsts)), "hipMalloc(film consts)")) return nullptr;
    } else if (check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize")) {
        return nullptr;  // a launch still reading the previous constants must finish first
    }
    if (check_hip(hipMemcpy(cc.dev, &h, sizeof(h), hipMemcpyHostToDevice), "hipMemcpy(film consts)")) return nullptr;
    cc.last = *p;
    cc.valid = true;
    return cc.dev;
}

template <typename L, bool HAS_PABS>
__global__ __launch_bounds__(128) void multiphysics_kernel(
    L *__restrict__ Te, L *__restrict__ Tl, L *__restrict__ TsT, L *__restrict__ TsR,
    L *__restrict__ mTx, L *__restrict__ mTy, L *__restrict__ mTz,
    L *__restrict__ mRx, L *__restrict__ mRy, L *__restrict__ mRz,
    double *__restrict__ window, const FilmConsts *__restrict__ gdp, const L *__restrict__ tm_lut,
    const L *__restrict__ re_lut, double T_min, double inv_dT, int32_t lut_N, double dt,
    int subcycles, int brillouin_iters, int64_t N) {
  const FilmConsts &gd = *gdp;
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

// Workgroups of the film update: at most 1024.  The kernel loops over the cells; on the
// 2.6 M-cell film a 1024-workgroup grid overlaps the Yee updates of the next step with 1%
// less step time than a full grid (512 was slower, 768 / 1280 / 2048 no faster).
static int64_t film_grid(int64_t N) {
    const int64_t full = cdiv(N, 128);
    return full < 1024 ? full : 1024;
}

template <typename L>
static int launch_mp(mp_film *f, const mp_ferri_params *p, const mp_lut *lut, double dt,
                     int subcycles, int iters, int has_pabs) {
    const int64_t N = f->n;
    // The launcher rounds dt, T_min and inv_dT to the compute type (Kernels.jl:662-663).
    const double dtc = (double)(L)dt, tmin = (double)(L)lut->T_min, invdt = (double)(L)lut->inv_dT;
    const FilmConsts *gc = film_consts(p);
    if (!gc) return MP_ERR_HIP;
    KernelTimer t("multiphysics");
#define MP_ARGS (L *)f->Te, (L *)f->Tl, (L *)f->Ts_TM, (L *)f->Ts_RE, (L *)f->m_TM_x, (L *)f->m_TM_y, \
                (L *)f->m_TM_z, (L *)f->m_RE_x, (L *)f->m_RE_y, (L *)f->m_RE_z, f->window_sum, gc,   \
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
