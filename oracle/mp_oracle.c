/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU restatement of MagnetoPhotonic.jl's coupled 3-D time step kernels (the reference
 * executes the same kernel bodies on its CPU backend through KernelAbstractions).  It is
 * the checker for libmphip: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  Its entry points mirror include/mphip.h with an `orc_`
 * prefix and take HOST pointers.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/fdtd/Kernels.jl:9-152    Yee H/E + CFS-CPML           -> orc_yee.h
 *   src/fdtd/Kernels.jl:154-355  sources, ADE, MO, pabs       -> orc_film.h
 *   src/fdtd/Kernels.jl:362-412  multiphysics (4TM + LLB)     -> orc_physics.h
 *   src/physics/Thermal.jl:24-40, src/physics/Magnetization.jl:7-263
 *   src/fdtd/Kernels.jl:420-478  DFT phase/accumulate, trace  -> below
 *   src/fdtd/Source.jl:188-236   plane-source host loop        -> below
 *   src/sim/Monitors.jl:263-341  FluxMonitor plane integral    -> below
 *   src/sim/ProbeReadout.jl:322-351  probe_shot plane reductions -> below
 *   src/sim/Monitors.jl:350-368, 461-473  mode overlap / plane mean -> below
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math -fopenmp (denormals kept).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include "../include/mphip.h"

#define ORC_EXPORT __attribute__((visibility("default")))

#include "orc_physics.h"
#include "orc_half.h"
#include "orc_surrogate.h"

#define PSI_LD(p, o) ((CT)(p)[o])
#define PSI_ST(p, o, v) ((p)[o] = (PT)(v))

/* ---- Yee instantiations: (ST, CT, PT) ------------------------------------------------- */
#define ST float
#define CT float
#define PT float
#define SFX(n) n##_fff
#define CFMA fmaf
#include "orc_yee.h"
#undef ST
#undef CT
#undef PT
#undef SFX
#undef CFMA

#define ST double
#define CT double
#define PT double
#define SFX(n) n##_ddd
#define CFMA fma
#include "orc_yee.h"
#undef ST
#undef CT
#undef PT
#undef SFX
#undef CFMA

#define ST float
#define CT double
#define PT float
#define SFX(n) n##_fdf
#define CFMA fma
#include "orc_yee.h"
#undef ST
#undef CT
#undef PT
#undef SFX
#undef CFMA

#define ST double
#define CT double
#define PT float
#define SFX(n) n##_ddf
#define CFMA fma
#include "orc_yee.h"
#undef ST
#undef CT
#undef PT
#undef SFX
#undef CFMA
#undef PSI_LD
#undef PSI_ST

/* FP32 fields and arithmetic with psi stored as BF16 / binary16 holding psi*psc */
#define ST float
#define CT float
#define PT uint16_t
#define CFMA fmaf
#define SFX(n) n##_ffb
#define PSI_LD(p, o) (orc_bf16_to_f32((p)[o]) * pisc)
#define PSI_ST(p, o, v) ((p)[o] = orc_f32_to_bf16((v) * psc))
#include "orc_yee.h"
#undef SFX
#undef PSI_LD
#undef PSI_ST
#define SFX(n) n##_ffh
#define PSI_LD(p, o) (orc_f16_to_f32((p)[o]) * pisc)
#define PSI_ST(p, o, v) ((p)[o] = orc_f32_to_f16((v) * psc))
#include "orc_yee.h"
#undef SFX
#undef PSI_LD
#undef PSI_ST
#undef ST
#undef CT
#undef PT
#undef CFMA

/* ---- film / source instantiations: (ST, CT) -------------------------------------------- */
#define ST float
#define CT float
#define SFX2(n) n##_ff
#define CFMA fmaf
#include "orc_film.h"
#undef ST
#undef CT
#undef SFX2
#undef CFMA

#define ST double
#define CT double
#define SFX2(n) n##_dd
#define CFMA fma
#include "orc_film.h"
#undef ST
#undef CT
#undef SFX2
#undef CFMA

#define ST float
#define CT double
#define SFX2(n) n##_fd
#define CFMA fma
#include "orc_film.h"
#undef ST
#undef CT
#undef SFX2
#undef CFMA

static _Thread_local char orc_err[256];
static int orc_fail(const char *msg) {
    strncpy(orc_err, msg, sizeof(orc_err) - 1);
    return MP_ERR_ARG;
}
ORC_EXPORT const char *orc_last_error(void) { return orc_err; }
ORC_EXPORT int orc_num_threads(void) { return omp_get_max_threads(); }
/* threads of the OpenMP loops from here on (n <= 0: unchanged) */
ORC_EXPORT void orc_set_num_threads(int n) { if (n > 0) omp_set_num_threads(n); }

/* element-wise storage conversions of the reduced kinds, exported for their own tests:
 * out[i] = bits of x[i]*sc, and back, y[i] = value(h[i])*isc */
ORC_EXPORT void orc_store_reduced(int32_t kind, const float *x, uint16_t *out, int64_t n, float sc) {
    for (int64_t i = 0; i < n; ++i) orc_sst(out, i, kind, x[i], sc);
}
ORC_EXPORT void orc_load_reduced(int32_t kind, const uint16_t *h, float *y, int64_t n, float isc) {
    for (int64_t i = 0; i < n; ++i) y[i] = (float)orc_sld(h, i, kind, isc);
}

/* storage/compute pair code */
static int pair_code(int32_t st, int32_t ct) {
    if (st == MP_F32 && ct == MP_F32) return 0;
    if (st == MP_F64 && ct == MP_F64) return 1;
    if (st == MP_F32 && ct == MP_F64) return 2;
    return -1;
}

/* storage scale of reduced-precision psi: psi is kept as psi*sc */
static void psi_scales(const mp_cpml *c, float *sc, float *isc) {
    double s = (c->psi_dtype == MP_BF16 || c->psi_dtype == MP_F16) ? c->psi_scale : 1.0;
    if (!(s > 0)) s = 1.0;
    *sc = (float)s;
    *isc = (float)(1.0 / s);
}

ORC_EXPORT int orc_update_H_3d(const mp_yee *g, mp_cpml *c, double dt, double mu0) {
    if (g->nx < 2 || g->ny < 2 || g->nz < 2) return 0;
    int pc = pair_code(g->field_dtype, g->compute_dtype);
    float sc, isc;
    psi_scales(c, &sc, &isc);
    if (pc == 0 && c->psi_dtype == MP_F32) {
        orc_yee_H_fff(g, c, (float)dt / (float)mu0, sc, isc);
    } else if (pc == 0 && c->psi_dtype == MP_BF16) {
        orc_yee_H_ffb(g, c, (float)dt / (float)mu0, sc, isc);
    } else if (pc == 0 && c->psi_dtype == MP_F16) {
        orc_yee_H_ffh(g, c, (float)dt / (float)mu0, sc, isc);
    } else if (pc == 1 && c->psi_dtype == MP_F64) {
        orc_yee_H_ddd(g, c, dt / mu0, sc, isc);
    } else if (pc == 2 && c->psi_dtype == MP_F32) {
        orc_yee_H_fdf(g, c, dt / mu0, sc, isc);
    } else if (pc == 1 && c->psi_dtype == MP_F32) {
        orc_yee_H_ddf(g, c, dt / mu0, sc, isc);
    } else {
        return orc_fail("orc_update_H_3d: unsupported dtype combination");
    }
    return 0;
}

ORC_EXPORT int orc_update_E_3d(const mp_yee *g, mp_cpml *c, double dt) {
    if (g->nx < 2 || g->ny < 2 || g->nz < 2) return 0;
    int pc = pair_code(g->field_dtype, g->compute_dtype);
    float sc, isc;
    psi_scales(c, &sc, &isc);
    if (pc == 0 && c->psi_dtype == MP_F32) {
        orc_yee_E_fff(g, c, (float)dt, sc, isc);
    } else if (pc == 0 && c->psi_dtype == MP_BF16) {
        orc_yee_E_ffb(g, c, (float)dt, sc, isc);
    } else if (pc == 0 && c->psi_dtype == MP_F16) {
        orc_yee_E_ffh(g, c, (float)dt, sc, isc);
    } else if (pc == 1 && c->psi_dtype == MP_F64) {
        orc_yee_E_ddd(g, c, dt, sc, isc);
    } else if (pc == 2 && c->psi_dtype == MP_F32) {
        orc_yee_E_fdf(g, c, dt, sc, isc);
    } else if (pc == 1 && c->psi_dtype == MP_F32) {
        orc_yee_E_ddf(g, c, dt, sc, isc);
    } else {
        return orc_fail("orc_update_E_3d: unsupported dtype combination");
    }
    return 0;
}

ORC_EXPORT int orc_inject_soft_3d(const mp_yee *g, int32_t component, int64_t i, int64_t j,
                                  int64_t k, double value) {
    if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz)
        return orc_fail("orc_inject_soft_3d: index out of range");
    const int64_t o = i + g->nx * (j + g->ny * k);
    switch (pair_code(g->field_dtype, g->compute_dtype)) {
    case 0: orc_inject_soft_ff(g, component, o, (float)value); break;
    case 1: orc_inject_soft_dd(g, component, o, value); break;
    case 2: orc_inject_soft_fd(g, component, o, value); break;
    default: return orc_fail("orc_inject_soft_3d: unsupported dtypes");
    }
    return 0;
}

ORC_EXPORT int orc_inject_mode_x_3d(const mp_yee *g, int32_t component, int64_t i,
                                    const void *profile, double value, double hcorr) {
    if (i < 0 || i >= g->nx) return orc_fail("orc_inject_mode_x_3d: index out of range");
    switch (pair_code(g->field_dtype, g->compute_dtype)) {
    case 0: orc_inject_mode_x_ff(g, component, i, (const float *)profile, (float)value, (float)hcorr); break;
    case 1: orc_inject_mode_x_dd(g, component, i, (const double *)profile, value, hcorr); break;
    case 2: orc_inject_mode_x_fd(g, component, i, (const double *)profile, value, hcorr); break;
    default: return orc_fail("orc_inject_mode_x_3d: unsupported dtypes");
    }
    return 0;
}

/* inject!(::FieldState, ::Grid3D, ::PlaneSource) and _inject_3d_cell! (Source.jl:188-236):
 * fields.E[i,j,k] += w*val promotes to Float64 and rounds once into the stored type;
 * D += Float64(w*val) / Float64(inv_eps). */
static void plane_cell(const mp_yee *g, int code, int64_t o, double dv) {
    int f64 = g->field_dtype == MP_F64;
#define ADD(ptr, inc)                                                          \
    do {                                                                       \
        if (f64) ((double *)(ptr))[o] = ((double *)(ptr))[o] + (inc);          \
        else ((float *)(ptr))[o] = (float)((double)((float *)(ptr))[o] + (inc)); \
    } while (0)
#define IE(ptr) (f64 ? ((const double *)(ptr))[o] : (double)((const float *)(ptr))[o])
    switch (code) {
    case 1: ADD(g->ex, dv); ADD(g->dx, dv / IE(g->inv_eps_x)); break;
    case 2: ADD(g->ey, dv); ADD(g->dy, dv / IE(g->inv_eps_y)); break;
    case 3: ADD(g->ez, dv); ADD(g->dz, dv / IE(g->inv_eps_z)); break;
    case 4: ADD(g->hx, dv); break;
    case 5: ADD(g->hy, dv); break;
    case 6: ADD(g->hz, dv); break;
    default: break;
    }
#undef ADD
#undef IE
}

ORC_EXPORT int orc_inject_plane_3d(const mp_yee *g, int32_t component, int32_t axis, int64_t i0,
                                   double w0, int64_t i1, double w1, double value) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    const int64_t n_ax = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (axis < 0 || axis > 2 || i0 < 0 || i0 >= n_ax || i1 >= n_ax)
        return orc_fail("orc_inject_plane_3d: bad axis or index");
    int64_t idxs[2] = {i0, i1};
    double ws[2] = {w0, w1};
    /* Julia loop order: for (i,wi) in weights, j, k (x axis) etc.; each cell touched by
     * exactly one weight, order irrelevant when i0 != i1. */
    for (int t = 0; t < 2; ++t) {
        if (idxs[t] < 0) continue;
        double dv = ws[t] * value;
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j)
                for (int64_t i = 0; i < nx; ++i) {
                    int64_t c = axis == 0 ? i : (axis == 1 ? j : k);
                    if (c != idxs[t]) continue;
                    plane_cell(g, component, i + nx * (j + ny * k), dv);
                }
    }
    return 0;
}

static int reduced_kind(int32_t k) { return k == MP_BF16 || k == MP_F16; }

/* P/J storage of an ADE list: kind, P scale and J scale.  Reduced kinds keep P*scale and J
 * in P units, J*(dt/2)*scale; CT storage uses no scale. */
static int ade_storage(const mp_ade *s, double dt, float *sc, float *isc, float *jsc, float *ijsc) {
    if (!reduced_kind(s->state_dtype)) {
        *sc = *isc = *jsc = *ijsc = 1.0f;
        return s->dtype;
    }
    const double p = s->state_scale > 0 ? s->state_scale : 1.0;
    const double j = p * dt / 2.0;
    *sc = (float)p;
    *isc = (float)(1.0 / p);
    *jsc = (float)j;
    *ijsc = (float)(1.0 / j);
    return s->state_dtype;
}

ORC_EXPORT int orc_patch_E_dispersive(void *E, mp_ade *s, int32_t field_dtype, double dt) {
    if (s->n <= 0 || s->npoles <= 0) return 0;
    float sc, isc, jsc, ijsc;
    const int sk = ade_storage(s, dt, &sc, &isc, &jsc, &ijsc);
    const int pc = pair_code(field_dtype, s->dtype);
    if (reduced_kind(sk) && pc != 0) return orc_fail("orc_patch_E_dispersive: reduced state needs FP32");
    switch (pc) {
    case 0: orc_ade_patch_ff((float *)E, s, 2.0f / (float)dt, sk, sc, isc, jsc, ijsc); break;
    case 1: orc_ade_patch_dd((double *)E, s, 2.0 / dt, sk, sc, isc, jsc, ijsc); break;
    case 2: orc_ade_patch_fd((float *)E, s, 2.0 / dt, sk, sc, isc, jsc, ijsc); break;
    default: return orc_fail("orc_patch_E_dispersive: unsupported dtypes");
    }
    return 0;
}

ORC_EXPORT int orc_patch_E_mo_gyration(void *Ea, void *Eb, mp_mo *s, const void *iea,
                                       const void *ieb, int32_t field_dtype, const void *m_TM,
                                       const void *m_RE, double Q_TM, double Q_RE, double dt) {
    if (s->n <= 0 || s->npoles <= 0) return 0;
    switch (pair_code(field_dtype, s->dtype)) {
    case 0:
        orc_mo_gyration_ff((float *)Ea, (float *)Eb, s, (const float *)iea, (const float *)ieb,
                           (const float *)m_TM, (const float *)m_RE, (float)Q_TM, (float)Q_RE,
                           2.0f / (float)dt);
        break;
    case 1:
        orc_mo_gyration_dd((double *)Ea, (double *)Eb, s, (const double *)iea, (const double *)ieb,
                           (const double *)m_TM, (const double *)m_RE, Q_TM, Q_RE, 2.0 / dt);
        break;
    case 2:
        orc_mo_gyration_fd((float *)Ea, (float *)Eb, s, (const float *)iea, (const float *)ieb,
                           (const double *)m_TM, (const double *)m_RE, Q_TM, Q_RE, 2.0 / dt);
        break;
    default: return orc_fail("orc_patch_E_mo_gyration: unsupported dtypes");
    }
    return 0;
}

ORC_EXPORT int orc_pabs_accumulate(mp_film *f, const mp_yee *g, const mp_ade *ax, const mp_ade *ay,
                                   const mp_ade *az, int32_t has_ade_work, double pabs_factor,
                                   double dt) {
    if (f->n <= 0) return 0;
    int ct = has_ade_work ? ax->dtype : f->dtype;
    float sc = 1.0f, isc = 1.0f, jsc = 1.0f, ijsc = 1.0f;
    const int sk = has_ade_work ? ade_storage(ax, dt, &sc, &isc, &jsc, &ijsc) : ct;
    switch (pair_code(g->field_dtype, ct)) {
    case 0: orc_pabs_ff(f, g, ax, ay, az, has_ade_work, pabs_factor, dt, sk, ijsc); break;
    case 1: orc_pabs_dd(f, g, ax, ay, az, has_ade_work, pabs_factor, dt, sk, ijsc); break;
    case 2: orc_pabs_fd(f, g, ax, ay, az, has_ade_work, pabs_factor, dt, sk, ijsc); break;
    default: return orc_fail("orc_pabs_accumulate: unsupported dtypes");
    }
    return 0;
}

/* mp_film_epilogue: the film list work after the E update, launched as the reference
 * launches it (_step_typed!, src/fdtd/Solver.jl:310-345): the ADE patch of Ex, Ey, Ez
 * (Kernels.jl:595-604), the gyration (:606-619), the absorbed power (:623-646). */
ORC_EXPORT int orc_film_epilogue(const mp_yee *g, mp_ade *ax, mp_ade *ay, mp_ade *az, mp_mo *mo,
                                 void *mo_ea, void *mo_eb, const void *mo_iea, const void *mo_ieb,
                                 mp_film *f, int32_t has_ade_work, double pabs_factor, double Q_TM,
                                 double Q_RE, double dt) {
    if (f->n <= 0) return 0;
    void *E[3] = {g->ex, g->ey, g->ez};
    mp_ade *a[3] = {ax, ay, az};
    for (int c = 0; c < 3; ++c) {
        const int rc = orc_patch_E_dispersive(E[c], a[c], g->field_dtype, dt);
        if (rc) return rc;
    }
    if (mo && mo->n > 0) {
        const int rc = orc_patch_E_mo_gyration(mo_ea, mo_eb, mo, mo_iea, mo_ieb, g->field_dtype, f->m_TM_x,
                                               f->m_RE_x, Q_TM, Q_RE, dt);
        if (rc) return rc;
    }
    return orc_pabs_accumulate(f, g, ax, ay, az, has_ade_work, pabs_factor, dt);
}

ORC_EXPORT int orc_multiphysics_step(mp_film *f, const mp_ferri_params *p, const mp_lut *lut,
                                     double dt, int32_t subcycles, int32_t brillouin_iters,
                                     int32_t has_pabs) {
    if (f->n <= 0) return 0;
    if (subcycles <= 0) return orc_fail("orc_multiphysics_step: subcycles must be positive");
    if (f->dtype == MP_F32) {
        orc_multiphysics_float(f, p, (const float *)lut->tm, (const float *)lut->re,
                               (double)(float)lut->T_min, (double)(float)lut->inv_dT, lut->n,
                               (double)(float)dt, subcycles, brillouin_iters, has_pabs);
    } else if (f->dtype == MP_F64) {
        orc_multiphysics_double(f, p, (const double *)lut->tm, (const double *)lut->re,
                                lut->T_min, lut->inv_dT, lut->n, dt, subcycles, brillouin_iters,
                                has_pabs);
    } else {
        return orc_fail("orc_multiphysics_step: unsupported dtype");
    }
    return 0;
}

/* Single-cell llb_step and update_4tm for the known-answer tests (Magnetization.jl:105,
 * Thermal.jl:33).  m[6] in/out, out[4] = dm_TM_x, dm_RE_x, Hani_TM_x, Hani_RE_x. */
ORC_EXPORT void orc_llb_step(double *m, double te, double tsT, double tsR, const mp_ferri_params *p,
                             const double *tm_lut, const double *re_lut, double T_min,
                             double inv_dT, int32_t lut_N, double dt, int32_t iters, double *out) {
    orc_llb_step_double(m, te, tsT, tsR, p, tm_lut, re_lut, T_min, inv_dT, lut_N, dt, iters, out);
}
ORC_EXPORT void orc_update_4tm_cell(double *T4, double pabs, double QT, double QR,
                                    const mp_ferri_params *p, double dt) {
    orc_update_4tm(&T4[0], &T4[1], &T4[2], &T4[3], pabs, QT, QR, p, dt);
}
ORC_EXPORT double orc_brillouin_fn(double S, double x) { return orc_brillouin(S, x); }
ORC_EXPORT double orc_lut_lookup(const double *lut, double T, double T_min, double inv_dT,
                                 int32_t n) {
    return orc_lut_lookup_double(lut, T, T_min, inv_dT, n);
}

/* ---- monitors ------------------------------------------------------------------------ */
static inline double load_f(const void *p, int32_t dt, int64_t o) {
    return dt == MP_F64 ? ((const double *)p)[o] : (double)((const float *)p)[o];
}

/* plane geometry: normal axis -> in-plane sizes and the tangential (a, b) components */
static void plane_geom(const mp_yee *g, int axis, int64_t *n1, int64_t *n2, const void **Ea,
                       const void **Eb, const void **Ha, const void **Hb) {
    if (axis == 0) {
        *n1 = g->ny; *n2 = g->nz;
        *Ea = g->ey; *Eb = g->ez; *Ha = g->hy; *Hb = g->hz;
    } else if (axis == 1) {
        *n1 = g->nz; *n2 = g->nx;
        *Ea = g->ez; *Eb = g->ex; *Ha = g->hz; *Hb = g->hx;
    } else {
        *n1 = g->nx; *n2 = g->ny;
        *Ea = g->ex; *Eb = g->ey; *Ha = g->hx; *Hb = g->hy;
    }
}

static inline int64_t plane_lin(const mp_yee *g, int axis, int64_t idx, int64_t a, int64_t b) {
    /* (a, b) are the in-plane coordinates in plane order */
    if (axis == 0) return idx + g->nx * (a + g->ny * b);        /* (j, k) */
    if (axis == 1) return b + g->nx * (idx + g->ny * a);        /* (k, i) */
    return a + g->nx * (b + g->ny * idx);                       /* (i, j) */
}

/* probe_phase_factors_kernel! + probe_dft_accumulate_kernel! (Kernels.jl:420-461) */
ORC_EXPORT int orc_probe_dft_accumulate(const mp_yee *g, const mp_plane *planes, int32_t nplanes,
                                        const double *omega_bins, double *phase, int32_t nb,
                                        double t, double dt) {
    for (int q = 0; q < nb; ++q) {
        double ph = omega_bins[q] * t;
        phase[2 * q] = cos(ph) * dt;
        phase[2 * q + 1] = -sin(ph) * dt;
    }
    for (int p = 0; p < nplanes; ++p) {
        const mp_plane *pl = &planes[p];
        int64_t n1, n2;
        const void *Ea, *Eb, *Ha, *Hb;
        plane_geom(g, pl->axis, &n1, &n2, &Ea, &Eb, &Ha, &Hb);
        const int64_t np = n1 * n2;
        for (int64_t b = 0; b < n2; ++b)
            for (int64_t a = 0; a < n1; ++a) {
                const int64_t o = plane_lin(g, pl->axis, pl->index, a, b);
                const int64_t c = a + n1 * b;
                double ea = load_f(Ea, g->field_dtype, o), eb = load_f(Eb, g->field_dtype, o);
                double ha = load_f(Ha, g->field_dtype, o), hb = load_f(Hb, g->field_dtype, o);
                pl->energy[c] += (ea * hb - eb * ha) * dt;
                for (int q = 0; q < nb; ++q) {
                    const double pr = phase[2 * q], pi = phase[2 * q + 1];
                    const int64_t z = 2 * (c + np * q);
                    pl->dft_ea[z] += ea * pr; pl->dft_ea[z + 1] += ea * pi;
                    pl->dft_eb[z] += eb * pr; pl->dft_eb[z + 1] += eb * pi;
                    pl->dft_ha[z] += ha * pr; pl->dft_ha[z + 1] += ha * pi;
                    pl->dft_hb[z] += hb * pr; pl->dft_hb[z + 1] += hb * pi;
                }
            }
    }
    return 0;
}

/* probe_trace_plane_kernel! + sum(dims=(1,2)) (Kernels.jl:467-478; ProbeReadout.jl:286-291) */
ORC_EXPORT int orc_probe_trace_plane(const mp_yee *g, int32_t axis, int64_t i_pre, int64_t i_post,
                                     const double *mode_w, const double *area, double *out) {
    int64_t n1, n2;
    const void *Ea, *Eb, *Ha, *Hb;
    plane_geom(g, axis, &n1, &n2, &Ea, &Eb, &Ha, &Hb);
    double s[4] = {0, 0, 0, 0};
    for (int64_t b = 0; b < n2; ++b)
        for (int64_t a = 0; a < n1; ++a) {
            const int64_t c = a + n1 * b;
            double w = mode_w[c] * area[c];
            int64_t o1 = plane_lin(g, axis, i_pre, a, b), o2 = plane_lin(g, axis, i_post, a, b);
            s[0] += load_f(Ea, g->field_dtype, o1) * w;
            s[1] += load_f(Eb, g->field_dtype, o1) * w;
            s[2] += load_f(Ea, g->field_dtype, o2) * w;
            s[3] += load_f(Eb, g->field_dtype, o2) * w;
        }
    memcpy(out, s, sizeof(s));
    return 0;
}

/* record!(::FluxMonitor) 3-D branch (Monitors.jl:302-339) */
ORC_EXPORT int orc_flux_plane(const mp_yee *g, int32_t axis, int64_t index, const double *area,
                              double *out, int64_t slot) {
    int64_t n1, n2;
    const void *Ea, *Eb, *Ha, *Hb;
    plane_geom(g, axis, &n1, &n2, &Ea, &Eb, &Ha, &Hb);
    double flux = 0.0;
    for (int64_t b = 0; b < n2; ++b)
        for (int64_t a = 0; a < n1; ++a) {
            const int64_t o = plane_lin(g, axis, index, a, b);
            flux += (load_f(Ea, g->field_dtype, o) * load_f(Hb, g->field_dtype, o) -
                     load_f(Eb, g->field_dtype, o) * load_f(Ha, g->field_dtype, o)) *
                    area[a + n1 * b];
        }
    out[slot] = flux;
    return 0;
}

/* probe_shot's plane reductions (ProbeReadout.jl:322-351): per-bin complex cross-power
 * _plane_poynting, _plane_net_flux and _jones_overlap (before the division by the mode
 * norm), each summed over the plane in column-major cell order as the reference loops. */
static void orc_poynting(const mp_plane *p, int64_t np, int q, const double *area, double *re, double *im) {
    double sr = 0.0, si = 0.0;
    for (int64_t c = 0; c < np; ++c) {
        const int64_t z = 2 * (c + np * q);
        const double ar = p->dft_ea[z], ai = p->dft_ea[z + 1], br = p->dft_eb[z], bi = p->dft_eb[z + 1];
        const double har = p->dft_ha[z], hai = p->dft_ha[z + 1], hbr = p->dft_hb[z], hbi = p->dft_hb[z + 1];
        /* (a * conj(hb) - b * conj(ha)) * dA */
        sr += ((ar * hbr + ai * hbi) - (br * har + bi * hai)) * area[c];
        si += ((ai * hbr - ar * hbi) - (bi * har - br * hai)) * area[c];
    }
    *re = sr;
    *im = si;
}

static void orc_overlap_bin(const double *dft, int64_t np, int q, const double *mode_w, const double *area,
                            double *out) {
    double sr = 0.0, si = 0.0;
    for (int64_t c = 0; c < np; ++c) {
        const double w = mode_w[c] * area[c];
        sr += dft[2 * (c + np * q)] * w;
        si += dft[2 * (c + np * q) + 1] * w;
    }
    out[0] = sr;
    out[1] = si;
}

ORC_EXPORT int orc_probe_shot_reduce(const mp_plane *pre, const mp_plane *post, int64_t n1, int64_t n2,
                                     int32_t nb, int32_t center_bin, const double *mode_w, const double *area,
                                     double *out) {
    const int64_t np = n1 * n2;
    for (int q = 0; q < nb; ++q) {
        orc_poynting(pre, np, q, area, &out[2 * q], &out[2 * q + 1]);
        orc_poynting(post, np, q, area, &out[2 * nb + 2 * q], &out[2 * nb + 2 * q + 1]);
    }
    double npre = 0.0, npost = 0.0;
    for (int64_t c = 0; c < np; ++c) {
        npre += pre->energy[c] * area[c];
        npost += post->energy[c] * area[c];
    }
    out[4 * nb] = npre;
    out[4 * nb + 1] = npost;
    orc_overlap_bin(post->dft_ea, np, center_bin, mode_w, area, &out[4 * nb + 2]);
    orc_overlap_bin(post->dft_eb, np, center_bin, mode_w, area, &out[4 * nb + 4]);
    orc_overlap_bin(pre->dft_ea, np, center_bin, mode_w, area, &out[4 * nb + 6]);
    orc_overlap_bin(pre->dft_eb, np, center_bin, mode_w, area, &out[4 * nb + 8]);
    return 0;
}

/* _mode_overlap (Monitors.jl:350-368) / the sum of _mean_plane_components
 * (Monitors.jl:461-473): sum(F * weight) over one plane, weight 1 when null. */
ORC_EXPORT int orc_plane_overlap(const mp_yee *g, int32_t component, int32_t axis, int64_t index,
                                 const double *weight, double *out) {
    int64_t n1, n2;
    const void *Ea, *Eb, *Ha, *Hb;
    plane_geom(g, axis, &n1, &n2, &Ea, &Eb, &Ha, &Hb);
    const void *comps[6] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz};
    const void *F = comps[component];
    double s = 0.0;
    for (int64_t b = 0; b < n2; ++b)
        for (int64_t a = 0; a < n1; ++a) {
            const double f = load_f(F, g->field_dtype, plane_lin(g, axis, index, a, b));
            s += weight ? f * weight[a + n1 * b] : f;
        }
    out[0] = s;
    return 0;
}

ORC_EXPORT int orc_film_reduce(const mp_film *f, double *out) {
    double s[8] = {0};
    double neg = 0.0, temax = -INFINITY;
    const void *arr[8] = {f->m_TM_x, f->m_TM_y, f->m_TM_z, f->m_RE_x, f->m_RE_y, f->m_RE_z, f->Te, f->Tl};
    for (int64_t n = 0; n < f->n; ++n) {
        for (int a = 0; a < 8; ++a) s[a] += load_f(arr[a], f->dtype, n);
        if (load_f(f->m_TM_x, f->dtype, n) < 0.0) neg += 1.0;
        double te = load_f(f->Te, f->dtype, n);
        if (te > temax) temax = te;
    }
    double inv = f->n > 0 ? 1.0 / (double)f->n : 0.0;
    for (int a = 0; a < 8; ++a) out[a] = s[a] * inv;
    out[8] = neg * inv;
    out[9] = f->n > 0 ? temax : 0.0;
    return 0;
}

/* The sums of one FilmAverage / SwitchedFraction record (Monitors.jl:487-497, 518-570;
 * reduce_mean, reduce_norm_mean, reduce_count_negative, Backend.jl:48-57). */
ORC_EXPORT int orc_film_average(const mp_film *f, double *out) {
    double s[13] = {0};
    for (int64_t n = 0; n < f->n; ++n) {
        const double tx = load_f(f->m_TM_x, f->dtype, n), ty = load_f(f->m_TM_y, f->dtype, n),
                     tz = load_f(f->m_TM_z, f->dtype, n);
        const double rx = load_f(f->m_RE_x, f->dtype, n), ry = load_f(f->m_RE_y, f->dtype, n),
                     rz = load_f(f->m_RE_z, f->dtype, n);
        s[0] += tx; s[1] += ty; s[2] += tz;
        s[3] += rx; s[4] += ry; s[5] += rz;
        s[6] += sqrt(tx * tx + ty * ty + tz * tz);
        s[7] += sqrt(rx * rx + ry * ry + rz * rz);
        s[8] += load_f(f->Te, f->dtype, n);
        s[9] += load_f(f->Tl, f->dtype, n);
        s[10] += load_f(f->Ts_TM, f->dtype, n);
        s[11] += load_f(f->Ts_RE, f->dtype, n);
        if (tx < 0.0) s[12] += 1.0;
    }
    memcpy(out, s, sizeof(s));
    return 0;
}

/* switching_metrics (Monitors.jl:682-727) as the counts and sums of mp_switching_metrics */
ORC_EXPORT int orc_switching_metrics(const mp_film *f, double core_fraction, double *out) {
    if (f->n <= 0) return orc_fail("orc_switching_metrics: empty film");
    double s[13] = {0};
    double umax = -INFINITY;
    int64_t imax = -1;
    for (int64_t n = 0; n < f->n; ++n) {
        const double tm = load_f(f->m_TM_x, f->dtype, n), re = load_f(f->m_RE_x, f->dtype, n);
        const int sw = tm < 0.0 && re > 0.0;
        s[0] += sw; s[1] += tm < 0.0; s[2] += re > 0.0;
        if (sw) { s[3] += tm; s[4] += re; } else { s[5] += tm; s[6] += re; }
        if (imax < 0 || f->U_abs[n] > umax) { umax = f->U_abs[n]; imax = n; }
    }
    s[7] = umax;
    s[8] = (double)imax;
    const double thr = core_fraction * umax;
    for (int64_t n = 0; n < f->n; ++n) {
        if (f->U_abs[n] >= thr) {
            const double tm = load_f(f->m_TM_x, f->dtype, n), re = load_f(f->m_RE_x, f->dtype, n);
            s[9] += 1.0;
            s[10] += (tm < 0.0 && re > 0.0);
            s[11] += tm;
            s[12] += re;
        }
    }
    memcpy(out, s, sizeof(s));
    return 0;
}

/* Training rows of the surrogate (restates mp_surrogate_dataset). */
ORC_EXPORT int orc_surrogate_dataset(const mp_film *f, const mp_ferri_params *p, const mp_lut *lut,
                                     double dt, int32_t subcycles, int32_t iters, int32_t has_pabs,
                                     int64_t stride, float *X, float *T, int64_t max_rows) {
    if (subcycles <= 0 || stride < 1) return orc_fail("orc_surrogate_dataset: bad arguments");
    int64_t rows = (f->n + stride - 1) / stride;
    if (rows > max_rows) rows = max_rows;
    if (rows <= 0) return 0;
    if (f->dtype == MP_F32)
        orc_dataset_float(f, p, (const float *)lut->tm, (const float *)lut->re, (double)(float)lut->T_min,
                          (double)(float)lut->inv_dT, lut->n, (double)(float)dt, subcycles, iters, has_pabs,
                          stride, rows, X, T);
    else
        orc_dataset_double(f, p, (const double *)lut->tm, (const double *)lut->re, lut->T_min, lut->inv_dT,
                           lut->n, dt, subcycles, iters, has_pabs, stride, rows, X, T);
    return 0;
}

/* Surrogate step given the network's raw rates (f->n rows of 6, host FP32); restates
 * the film side of mp_surrogate_step.  *fallbacks receives the fallback count. */
ORC_EXPORT int orc_surrogate_apply(mp_film *f, const mp_ferri_params *p, const mp_lut *lut,
                                   const float *rates, double dt, int32_t subcycles, int32_t iters,
                                   int32_t has_pabs, double tol, int64_t *fallbacks) {
    if (subcycles <= 0) return orc_fail("orc_surrogate_apply: subcycles must be positive");
    int64_t fb = 0;
    if (f->n > 0) {
        if (f->dtype == MP_F32)
            fb = orc_surrogate_float(f, p, (const float *)lut->tm, (const float *)lut->re,
                                     (double)(float)lut->T_min, (double)(float)lut->inv_dT, lut->n, rates,
                                     (double)(float)dt, subcycles, iters, has_pabs, tol);
        else
            fb = orc_surrogate_double(f, p, (const double *)lut->tm, (const double *)lut->re, lut->T_min,
                                      lut->inv_dT, lut->n, rates, dt, subcycles, iters, has_pabs, tol);
    }
    if (fallbacks) *fallbacks = fb;
    return 0;
}
