// List kernels over the dispersive / film cells: Drude-Lorentz ADE patch, magneto-optic
// gyration and absorbed-power accumulation.
//
// Reference: ade_patch_kernel! (src/fdtd/Kernels.jl:194-231), mo_gyration_kernel!
// (:233-305), pabs_accumulate_kernel! (:313-355) and their launchers (:595-646).
// One thread per list entry; P/J are (n, npoles) column-major so a wave reads 64
// consecutive entries of one pole.  E is gathered through the cell index list.
#include "common.h"

namespace mp {

// P/J storage: CT, or BF16/F16 holding value*scale (FP32 arithmetic).  enc/dec are the
// conversions of one value; dec(enc(v)) is the value a later load of the stored v gives.
template <typename C, typename T> struct StateIO {
    static __device__ __forceinline__ T enc(C v, float) { return (T)v; }
    static __device__ __forceinline__ C dec(T q, float) { return (C)q; }
    static __device__ __forceinline__ C load(const T *p, int64_t o, float isc) { return dec(p[o], isc); }
    static __device__ __forceinline__ void store(T *p, int64_t o, C v, float sc) { p[o] = enc(v, sc); }
};
template <typename T> struct StateIOReduced {
    static __device__ __forceinline__ T enc(float v, float sc) { return st<T>(v * sc); }
    static __device__ __forceinline__ float dec(T q, float isc) { return ld<float>(q) * isc; }
    static __device__ __forceinline__ float load(const T *p, int64_t o, float isc) { return dec(p[o], isc); }
    static __device__ __forceinline__ void store(T *p, int64_t o, float v, float sc) { p[o] = enc(v, sc); }
};
template <> struct StateIO<float, bf16> : StateIOReduced<bf16> {};
template <> struct StateIO<float, f16s> : StateIOReduced<f16s> {};

constexpr int MAXP = 8;

template <typename S, typename C, typename T>
__global__ __launch_bounds__(256) void ade_patch_kernel(
    S *__restrict__ E, T *__restrict__ P, T *__restrict__ J, C *__restrict__ E_old,
    const int32_t *__restrict__ idxs, const C *__restrict__ fill, const C *__restrict__ inv_eps,
    const C *__restrict__ poles, int np, int64_t N, C inv_alpha, float sc, float isc, float jsc,
    float ijsc) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t idx = idxs[n];
    const C Eprev = E_old[n];
    const C f = fill[n];
    const C ie = inv_eps[n];
    C Pold[MAXP], Jold[MAXP];
    C sum_psi = (C)0, sum_beta = (C)0;
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
        if (p >= np) break;
        Pold[p] = StateIO<C, T>::load(P, n + p * N, isc);
        Jold[p] = StateIO<C, T>::load(J, n + p * N, ijsc);
        const C beta = poles[3 * p + 2] * f;
        const C psi = fmaT<C>(poles[3 * p], Pold[p], fmaT<C>(poles[3 * p + 1], Jold[p], beta * Eprev));
        sum_psi += psi;
        sum_beta += beta;
    }
    const C den = (C)1 + sum_beta * ie;
    const C Etemp = (C)E[idx];
    if (finiteT(Etemp) && finiteT(sum_psi) && finiteT(den) && __builtin_fabs((double)den) >= (double)(C)1e-30) {
        const C Efinal = (Etemp - sum_psi * ie) / den;
        E[idx] = (S)Efinal;
        E_old[n] = Efinal;
#pragma unroll
        for (int p = 0; p < MAXP; ++p) {
            if (p >= np) break;
            const C beta = poles[3 * p + 2] * f;
            const C psi = fmaT<C>(poles[3 * p], Pold[p], fmaT<C>(poles[3 * p + 1], Jold[p], beta * Eprev));
            const C Pnew = psi + beta * Efinal;
            StateIO<C, T>::store(P, n + p * N, Pnew, sc);
            StateIO<C, T>::store(J, n + p * N, (Pnew - Pold[p]) * inv_alpha - Jold[p], jsc);
        }
    }
}

template <typename S, typename C>
__global__ __launch_bounds__(256) void mo_gyration_kernel(
    S *__restrict__ Ea, S *__restrict__ Eb, C *__restrict__ Pa, C *__restrict__ Ja,
    C *__restrict__ Pb, C *__restrict__ Jb, C *__restrict__ Eoa, C *__restrict__ Eob,
    const int32_t *__restrict__ idxs, const C *__restrict__ fill, const int32_t *__restrict__ mposv,
    const S *__restrict__ iea, const S *__restrict__ ieb, const C *__restrict__ m_TM,
    const C *__restrict__ m_RE, const C *__restrict__ poles, int np, int64_t N, C Q_TM, C Q_RE,
    C inv_alpha) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t idx = idxs[n];
    const int32_t mpos = mposv[n];
    if (mpos < 0) return;
    const C f = fill[n];
    const C qeff = Q_TM * m_TM[mpos] + Q_RE * m_RE[mpos];
    const C Ea_drive = (C)Ea[idx];
    const C Eb_drive = (C)Eb[idx];
    const C Ea_old = Eoa[n];
    const C Eb_old = Eob[n];
    C sum_psi_a = (C)0, sum_psi_b = (C)0, sumC3 = (C)0;
    for (int p = 0; p < np; ++p) {
        const C C1 = poles[3 * p], C2 = poles[3 * p + 1], C3 = poles[3 * p + 2];
        const C beta = C3 * f * qeff;
        sum_psi_a += fmaT<C>(C1, Pa[n + p * N], fmaT<C>(C2, Ja[n + p * N], beta * Eb_old));
        sum_psi_b += fmaT<C>(C1, Pb[n + p * N], fmaT<C>(C2, Jb[n + p * N], beta * Ea_old));
        sumC3 += C3;
    }
    const C ica = (C)iea[idx];
    const C icb = (C)ieb[idx];
    const C Bmo = f * qeff * sumC3;
    const C half = (C)0.5;
    const C aa = half * ica * Bmo;
    const C ab = half * icb * Bmo;
    const C RHSa = Ea_drive - ica * sum_psi_a - aa * Eb_drive;
    const C RHSb = Eb_drive + icb * sum_psi_b + ab * Ea_drive;
    const C den = (C)1 + aa * ab;
    const C Eb_final = (RHSb + ab * RHSa) / den;
    const C Ea_final = RHSa - aa * Eb_final;
    if (finiteT(Ea_final) && finiteT(Eb_final)) {
        Ea[idx] = (S)Ea_final;
        Eb[idx] = (S)Eb_final;
        Eoa[n] = Ea_final;
        Eob[n] = Eb_final;
        const C Ea_mid = half * (Ea_drive + Ea_final);
        const C Eb_mid = half * (Eb_drive + Eb_final);
        for (int p = 0; p < np; ++p) {
            const C C1 = poles[3 * p], C2 = poles[3 * p + 1], C3 = poles[3 * p + 2];
            const C beta = C3 * f * qeff;
            const C Pa_old = Pa[n + p * N], Ja_old = Ja[n + p * N];
            const C Pa_new = fmaT<C>(C1, Pa_old, fmaT<C>(C2, Ja_old, beta * Eb_old)) + beta * Eb_mid;
            Pa[n + p * N] = Pa_new;
            Ja[n + p * N] = (Pa_new - Pa_old) * inv_alpha - Ja_old;
            const C Pb_old = Pb[n + p * N], Jb_old = Jb[n + p * N];
            const C Pb_new = fmaT<C>(C1, Pb_old, fmaT<C>(C2, Jb_old, beta * Ea_old)) + beta * Ea_mid;
            Pb[n + p * N] = Pb_new;
            Jb[n + p * N] = (Pb_new - Pb_old) * inv_alpha - Jb_old;
        }
    }
}

template <typename S, typename C, typename T, bool ADE_WORK>
__global__ __launch_bounds__(256) void pabs_kernel(
    double *__restrict__ U_abs, double *__restrict__ window, double *__restrict__ peak,
    const S *__restrict__ Ex, const S *__restrict__ Ey, const S *__restrict__ Ez,
    const int32_t *__restrict__ cells, const int32_t *__restrict__ pos_x,
    const int32_t *__restrict__ pos_y, const int32_t *__restrict__ pos_z,
    const T *__restrict__ Jx, const T *__restrict__ Jy, const T *__restrict__ Jz,
    int64_t nJx, int64_t nJy, int64_t nJz, int np, double pabs_factor, double dt, int64_t N,
    float isc) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t li = cells[n];
    const double ex = (double)Ex[li], ey = (double)Ey[li], ez = (double)Ez[li];
    double pabs;
    if (ADE_WORK) {
        double jx = 0.0, jy = 0.0, jz = 0.0;
        const int32_t px = pos_x[n], py = pos_y[n], pz = pos_z[n];
        for (int p = 0; p < np; ++p) {
            if (px >= 0) jx += (double)StateIO<C, T>::load(Jx, px + p * nJx, isc);
            if (py >= 0) jy += (double)StateIO<C, T>::load(Jy, py + p * nJy, isc);
            if (pz >= 0) jz += (double)StateIO<C, T>::load(Jz, pz + p * nJz, isc);
        }
        pabs = jl_max(ex * jx + ey * jy + ez * jz, 0.0);
    } else {
        pabs = pabs_factor * (ex * ex + ey * ey + ez * ez);
    }
    const double pdt = pabs * dt;
    U_abs[n] += pdt;
    window[n] += pdt;
    if (pabs > peak[n]) peak[n] = pabs;
}

// Reduced-precision P/J storage (BF16, FP16): P is stored as P*scale and J in P units,
// J*(dt/2)*scale, so one scale fits both (|J| ~ w|P| is ~1e15 larger than |P|).
static inline bool reduced(int32_t d) { return d == MP_BF16 || d == MP_F16; }
static inline double state_scale(const mp_ade *s) {
    return reduced(s->state_dtype) && s->state_scale > 0 ? s->state_scale : 1.0;
}

template <typename S, typename C, typename T>
static int launch_ade(void *E, mp_ade *s, double dt) {
    const int64_t N = s->n;
    const double sc = state_scale(s);
    const C inv_alpha = (C)2 / (C)dt;
    const double jsc = reduced(s->state_dtype) ? sc * dt / 2.0 : 1.0;
    KernelTimer t("ade_patch");
    ade_patch_kernel<S, C, T><<<cdiv(N, 256), 256, 0, stream()>>>(
        (S *)E, (T *)s->P, (T *)s->J, (C *)s->E_old, s->idx, (const C *)s->fill,
        (const C *)s->inv_eps, (const C *)s->poles, s->npoles, N, inv_alpha, (float)sc,
        (float)(1.0 / sc), (float)jsc, (float)(1.0 / jsc));
    MP_LAUNCH_CHECK("ade_patch_kernel");
    return MP_OK;
}

template <typename S, typename C, typename T>
static int launch_pabs(mp_film *f, const mp_yee *g, const mp_ade *ax, const mp_ade *ay,
                       const mp_ade *az, int has_ade_work, double pabs_factor, double dt) {
    const int64_t N = f->n;
    KernelTimer t("pabs");
    if (has_ade_work) {
        // J load factor: 1 for CT storage, 1/(scale*dt/2) for reduced storage
        const double sc = reduced(ax->state_dtype) ? state_scale(ax) * dt / 2.0 : 1.0;
        pabs_kernel<S, C, T, true><<<cdiv(N, 256), 256, 0, stream()>>>(
            f->U_abs, f->window_sum, f->peak, (const S *)g->ex, (const S *)g->ey, (const S *)g->ez,
            f->cells, f->pos_x, f->pos_y, f->pos_z, (const T *)ax->J, (const T *)ay->J,
            (const T *)az->J, ax->n, ay->n, az->n, ax->npoles, pabs_factor, dt, N, (float)(1.0 / sc));
    } else {
        pabs_kernel<S, C, T, false><<<cdiv(N, 256), 256, 0, stream()>>>(
            f->U_abs, f->window_sum, f->peak, (const S *)g->ex, (const S *)g->ey, (const S *)g->ez,
            f->cells, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, 0,
            pabs_factor, dt, N, 1.0f);
    }
    MP_LAUNCH_CHECK("pabs_kernel");
    return MP_OK;
}


// ---- fused film epilogue ------------------------------------------------------------------
// The step's list work after the E update in one pass over the film cells: the ADE patch of
// Ex, Ey and Ez, then the gyrotropic exchange of the pair (a, b), then the absorbed power.
// Every list is the film list (entry n of each ADE list, of the MO list and of the film is
// the same cell), so each cell runs exactly the operations of ade_patch_kernel x3,
// mo_gyration_kernel and pabs_kernel, in that order, on its own values: the result is that
// of the separate kernels bit for bit.  E is read and written once per cell instead of by
// every kernel, J stays in registers for the absorbed power, and no position map is read.
template <typename C, typename T> struct AdeView {
    T *P, *J;
    C *E_old;
    const C *fill, *inv_eps;
    float sc, isc, jsc, ijsc;   // P and J storage scales and their inverses (1 for CT storage)
};
template <typename C> struct MoView {
    C *Pa, *Ja, *Pb, *Jb, *Eoa, *Eob;
    const C *fill, *m_TM, *m_RE;
};

// ade_patch_kernel on one entry whose E is held in `Ecell` (storage type S).  jsum is the sum
// over poles of J as stored after the call (decoded as pabs_kernel loads it), in the order
// pabs_kernel adds it.
template <typename S, typename C, typename T>
__device__ __forceinline__ void ade_entry(S &Ecell, const AdeView<C, T> &a, const C *__restrict__ poles, int np,
                                          int64_t n, int64_t N, C inv_alpha, double &jsum) {
    const C Eprev = a.E_old[n];
    const C f = a.fill[n];
    const C ie = a.inv_eps[n];
    C Pold[MAXP], Jold[MAXP];
    C sum_psi = (C)0, sum_beta = (C)0;
#pragma unroll
    for (int p = 0; p < MAXP; ++p) {
        if (p >= np) break;
        Pold[p] = StateIO<C, T>::load(a.P, n + p * N, a.isc);
        Jold[p] = StateIO<C, T>::load(a.J, n + p * N, a.ijsc);
        const C beta = poles[3 * p + 2] * f;
        const C psi = fmaT<C>(poles[3 * p], Pold[p], fmaT<C>(poles[3 * p + 1], Jold[p], beta * Eprev));
        sum_psi += psi;
        sum_beta += beta;
    }
    const C den = (C)1 + sum_beta * ie;
    const C Etemp = (C)Ecell;
    jsum = 0.0;
    if (finiteT(Etemp) && finiteT(sum_psi) && finiteT(den) && __builtin_fabs((double)den) >= (double)(C)1e-30) {
        const C Efinal = (Etemp - sum_psi * ie) / den;
        Ecell = (S)Efinal;
        a.E_old[n] = Efinal;
#pragma unroll
        for (int p = 0; p < MAXP; ++p) {
            if (p >= np) break;
            const C beta = poles[3 * p + 2] * f;
            const C psi = fmaT<C>(poles[3 * p], Pold[p], fmaT<C>(poles[3 * p + 1], Jold[p], beta * Eprev));
            const C Pnew = psi + beta * Efinal;
            const C Jnew = (Pnew - Pold[p]) * inv_alpha - Jold[p];
            StateIO<C, T>::store(a.P, n + p * N, Pnew, a.sc);
            const T jq = StateIO<C, T>::enc(Jnew, a.jsc);
            a.J[n + p * N] = jq;
            jsum += (double)StateIO<C, T>::dec(jq, a.ijsc);
        }
    } else {
#pragma unroll
        for (int p = 0; p < MAXP; ++p) {
            if (p >= np) break;
            jsum += (double)Jold[p];
        }
    }
}

template <typename S, typename C, typename T, bool HAS_MO, bool ADE_WORK>
__global__ __launch_bounds__(256) void film_epilogue_kernel(
    S *__restrict__ Ex, S *__restrict__ Ey, S *__restrict__ Ez, const int32_t *__restrict__ cells,
    AdeView<C, T> ax, AdeView<C, T> ay, AdeView<C, T> az, const C *__restrict__ poles, int np, C inv_alpha,
    MoView<C> mo, const C *__restrict__ mo_poles, int mo_np, int mo_a, int mo_b,
    const S *__restrict__ iea, const S *__restrict__ ieb, C Q_TM, C Q_RE, C mo_inv_alpha,
    double *__restrict__ U_abs, double *__restrict__ window, double *__restrict__ peak, double pabs_factor,
    double dt, int64_t N) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int64_t idx = cells[n];
    S e[3] = {Ex[idx], Ey[idx], Ez[idx]};
    double js[3];
    ade_entry<S, C, T>(e[0], ax, poles, np, n, N, inv_alpha, js[0]);
    ade_entry<S, C, T>(e[1], ay, poles, np, n, N, inv_alpha, js[1]);
    ade_entry<S, C, T>(e[2], az, poles, np, n, N, inv_alpha, js[2]);
    if (HAS_MO) {
        // mo_gyration_kernel on entry n (its film position is n)
        const C f = mo.fill[n];
        const C qeff = Q_TM * mo.m_TM[n] + Q_RE * mo.m_RE[n];
        const S ea_s = mo_a == 0 ? e[0] : (mo_a == 1 ? e[1] : e[2]);
        const S eb_s = mo_b == 0 ? e[0] : (mo_b == 1 ? e[1] : e[2]);
        const C Ea_drive = (C)ea_s;
        const C Eb_drive = (C)eb_s;
        const C Ea_old = mo.Eoa[n];
        const C Eb_old = mo.Eob[n];
        C sum_psi_a = (C)0, sum_psi_b = (C)0, sumC3 = (C)0;
        for (int p = 0; p < mo_np; ++p) {
            const C C1 = mo_poles[3 * p], C2 = mo_poles[3 * p + 1], C3 = mo_poles[3 * p + 2];
            const C beta = C3 * f * qeff;
            sum_psi_a += fmaT<C>(C1, mo.Pa[n + p * N], fmaT<C>(C2, mo.Ja[n + p * N], beta * Eb_old));
            sum_psi_b += fmaT<C>(C1, mo.Pb[n + p * N], fmaT<C>(C2, mo.Jb[n + p * N], beta * Ea_old));
            sumC3 += C3;
        }
        const C ica = (C)iea[idx];
        const C icb = (C)ieb[idx];
        const C Bmo = f * qeff * sumC3;
        const C half = (C)0.5;
        const C aa = half * ica * Bmo;
        const C ab = half * icb * Bmo;
        const C RHSa = Ea_drive - ica * sum_psi_a - aa * Eb_drive;
        const C RHSb = Eb_drive + icb * sum_psi_b + ab * Ea_drive;
        const C den = (C)1 + aa * ab;
        const C Eb_final = (RHSb + ab * RHSa) / den;
        const C Ea_final = RHSa - aa * Eb_final;
        if (finiteT(Ea_final) && finiteT(Eb_final)) {
            const S ea_new = (S)Ea_final, eb_new = (S)Eb_final;
            if (mo_a == 0) e[0] = ea_new; else if (mo_a == 1) e[1] = ea_new; else e[2] = ea_new;
            if (mo_b == 0) e[0] = eb_new; else if (mo_b == 1) e[1] = eb_new; else e[2] = eb_new;
            mo.Eoa[n] = Ea_final;
            mo.Eob[n] = Eb_final;
            const C Ea_mid = half * (Ea_drive + Ea_final);
            const C Eb_mid = half * (Eb_drive + Eb_final);
            for (int p = 0; p < mo_np; ++p) {
                const C C1 = mo_poles[3 * p], C2 = mo_poles[3 * p + 1], C3 = mo_poles[3 * p + 2];
                const C beta = C3 * f * qeff;
                const C Pa_old = mo.Pa[n + p * N], Ja_old = mo.Ja[n + p * N];
                const C Pa_new = fmaT<C>(C1, Pa_old, fmaT<C>(C2, Ja_old, beta * Eb_old)) + beta * Eb_mid;
                mo.Pa[n + p * N] = Pa_new;
                mo.Ja[n + p * N] = (Pa_new - Pa_old) * mo_inv_alpha - Ja_old;
                const C Pb_old = mo.Pb[n + p * N], Jb_old = mo.Jb[n + p * N];
                const C Pb_new = fmaT<C>(C1, Pb_old, fmaT<C>(C2, Jb_old, beta * Ea_old)) + beta * Ea_mid;
                mo.Pb[n + p * N] = Pb_new;
                mo.Jb[n + p * N] = (Pb_new - Pb_old) * mo_inv_alpha - Jb_old;
            }
        }
    }
    Ex[idx] = e[0];
    Ey[idx] = e[1];
    Ez[idx] = e[2];
    // pabs_kernel on film cell n
    const double ex = (double)e[0], ey = (double)e[1], ez = (double)e[2];
    const double pabs = ADE_WORK ? jl_max(ex * js[0] + ey * js[1] + ez * js[2], 0.0)
                                 : pabs_factor * (ex * ex + ey * ey + ez * ez);
    const double pdt = pabs * dt;
    U_abs[n] += pdt;
    window[n] += pdt;
    if (pabs > peak[n]) peak[n] = pabs;
}

// The same per-cell operations with NP poles in every ADE list and in the MO list (the film's
// Drude + Lorentz ADE poles and its two probe MO poles: NP = 2), every load of the cell issued
// before its first store.  In film_epilogue_kernel a store through one list's arrays might
// alias a later load through another list's, so each list's loads wait behind the previous
// list's stores: one memory round trip per list.  The arithmetic and its order are unchanged.
// Every array is addressed by a 32-bit byte offset from its base in scalar registers (the
// launcher checks that the offsets fit), so the cell's loads share a few offset registers.
template <typename T> __device__ __forceinline__ T &bref(T *base, uint32_t boff) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff);
}
template <typename T> __device__ __forceinline__ const T &bref(const T *base, uint32_t boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
template <typename C, int NP> struct AdeCell {
    C Eprev, f, ie, P[NP], J[NP];
};
template <typename C, int NP> struct MoCell {
    C f, m_TM, m_RE, Eoa, Eob, Pa[NP], Ja[NP], Pb[NP], Jb[NP];
};
// oc: byte offset of entry n in a C array; ot: of entry n in a T array; stT: of one pole's
// block (N entries) in a T array
template <typename C, typename T, int NP>
__device__ __forceinline__ AdeCell<C, NP> ade_cell_load(const AdeView<C, T> &a, uint32_t oc, uint32_t ot,
                                                        uint32_t stT) {
    AdeCell<C, NP> c;
    c.Eprev = bref(a.E_old, oc);
    c.f = bref(a.fill, oc);
    c.ie = bref(a.inv_eps, oc);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        c.P[p] = StateIO<C, T>::dec(bref(a.P, ot + p * stT), a.isc);
        c.J[p] = StateIO<C, T>::dec(bref(a.J, ot + p * stT), a.ijsc);
    }
    return c;
}
// ade_entry's arithmetic on loaded values
template <typename S, typename C, typename T, int NP>
__device__ __forceinline__ void ade_cell_apply(S &Ecell, const AdeView<C, T> &a, const AdeCell<C, NP> &c,
                                               const C *__restrict__ poles, uint32_t oc, uint32_t ot,
                                               uint32_t stT, C inv_alpha, double &jsum) {
    C sum_psi = (C)0, sum_beta = (C)0;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const C beta = poles[3 * p + 2] * c.f;
        const C psi = fmaT<C>(poles[3 * p], c.P[p], fmaT<C>(poles[3 * p + 1], c.J[p], beta * c.Eprev));
        sum_psi += psi;
        sum_beta += beta;
    }
    const C den = (C)1 + sum_beta * c.ie;
    const C Etemp = (C)Ecell;
    jsum = 0.0;
    if (finiteT(Etemp) && finiteT(sum_psi) && finiteT(den) && __builtin_fabs((double)den) >= (double)(C)1e-30) {
        const C Efinal = (Etemp - sum_psi * c.ie) / den;
        Ecell = (S)Efinal;
        bref(a.E_old, oc) = Efinal;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const C beta = poles[3 * p + 2] * c.f;
            const C psi = fmaT<C>(poles[3 * p], c.P[p], fmaT<C>(poles[3 * p + 1], c.J[p], beta * c.Eprev));
            const C Pnew = psi + beta * Efinal;
            const C Jnew = (Pnew - c.P[p]) * inv_alpha - c.J[p];
            bref(a.P, ot + p * stT) = StateIO<C, T>::enc(Pnew, a.sc);
            const T jq = StateIO<C, T>::enc(Jnew, a.jsc);
            bref(a.J, ot + p * stT) = jq;
            jsum += (double)StateIO<C, T>::dec(jq, a.ijsc);
        }
    } else {
#pragma unroll
        for (int p = 0; p < NP; ++p) jsum += (double)c.J[p];
    }
}

template <typename S, typename C, typename T, bool HAS_MO, bool ADE_WORK, int NP>
__global__ __launch_bounds__(256) void film_epilogue_np_kernel(
    S *__restrict__ Ex, S *__restrict__ Ey, S *__restrict__ Ez, const int32_t *__restrict__ cells,
    AdeView<C, T> ax, AdeView<C, T> ay, AdeView<C, T> az, const C *__restrict__ poles, int, C inv_alpha,
    MoView<C> mo, const C *__restrict__ mo_poles, int, int mo_a, int mo_b,
    const S *__restrict__ iea, const S *__restrict__ ieb, C Q_TM, C Q_RE, C mo_inv_alpha,
    double *__restrict__ U_abs, double *__restrict__ window, double *__restrict__ peak, double pabs_factor,
    double dt, int64_t N) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (uint64_t)N) return;
    const uint32_t oc = n * (uint32_t)sizeof(C), ot = n * (uint32_t)sizeof(T), od = n * 8u;
    const uint32_t stC = (uint32_t)N * (uint32_t)sizeof(C), stT = (uint32_t)N * (uint32_t)sizeof(T);
    // every load of the cell
    const uint32_t oe = (uint32_t)bref(cells, n * 4u) * (uint32_t)sizeof(S);
    S e[3] = {bref(Ex, oe), bref(Ey, oe), bref(Ez, oe)};
    const AdeCell<C, NP> cx = ade_cell_load<C, T, NP>(ax, oc, ot, stT);
    const AdeCell<C, NP> cy = ade_cell_load<C, T, NP>(ay, oc, ot, stT);
    const AdeCell<C, NP> cz = ade_cell_load<C, T, NP>(az, oc, ot, stT);
    MoCell<C, NP> m{};
    C ica = (C)0, icb = (C)0;
    if (HAS_MO) {
        m.f = bref(mo.fill, oc);
        m.m_TM = bref(mo.m_TM, oc);
        m.m_RE = bref(mo.m_RE, oc);
        m.Eoa = bref(mo.Eoa, oc);
        m.Eob = bref(mo.Eob, oc);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            m.Pa[p] = bref(mo.Pa, oc + p * stC);
            m.Ja[p] = bref(mo.Ja, oc + p * stC);
            m.Pb[p] = bref(mo.Pb, oc + p * stC);
            m.Jb[p] = bref(mo.Jb, oc + p * stC);
        }
        ica = (C)bref(iea, oe);
        icb = (C)bref(ieb, oe);
    }
    const double U_old = bref(U_abs, od), W_old = bref(window, od), peak_old = bref(peak, od);
    // ade_patch_kernel x3
    double js[3];
    ade_cell_apply<S, C, T, NP>(e[0], ax, cx, poles, oc, ot, stT, inv_alpha, js[0]);
    ade_cell_apply<S, C, T, NP>(e[1], ay, cy, poles, oc, ot, stT, inv_alpha, js[1]);
    ade_cell_apply<S, C, T, NP>(e[2], az, cz, poles, oc, ot, stT, inv_alpha, js[2]);
    if (HAS_MO) {
        // mo_gyration_kernel on entry n
        const C f = m.f;
        const C qeff = Q_TM * m.m_TM + Q_RE * m.m_RE;
        const S ea_s = mo_a == 0 ? e[0] : (mo_a == 1 ? e[1] : e[2]);
        const S eb_s = mo_b == 0 ? e[0] : (mo_b == 1 ? e[1] : e[2]);
        const C Ea_drive = (C)ea_s;
        const C Eb_drive = (C)eb_s;
        const C Ea_old = m.Eoa;
        const C Eb_old = m.Eob;
        C sum_psi_a = (C)0, sum_psi_b = (C)0, sumC3 = (C)0;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const C C1 = mo_poles[3 * p], C2 = mo_poles[3 * p + 1], C3 = mo_poles[3 * p + 2];
            const C beta = C3 * f * qeff;
            sum_psi_a += fmaT<C>(C1, m.Pa[p], fmaT<C>(C2, m.Ja[p], beta * Eb_old));
            sum_psi_b += fmaT<C>(C1, m.Pb[p], fmaT<C>(C2, m.Jb[p], beta * Ea_old));
            sumC3 += C3;
        }
        const C Bmo = f * qeff * sumC3;
        const C half = (C)0.5;
        const C aa = half * ica * Bmo;
        const C ab = half * icb * Bmo;
        const C RHSa = Ea_drive - ica * sum_psi_a - aa * Eb_drive;
        const C RHSb = Eb_drive + icb * sum_psi_b + ab * Ea_drive;
        const C den = (C)1 + aa * ab;
        const C Eb_final = (RHSb + ab * RHSa) / den;
        const C Ea_final = RHSa - aa * Eb_final;
        if (finiteT(Ea_final) && finiteT(Eb_final)) {
            const S ea_new = (S)Ea_final, eb_new = (S)Eb_final;
            if (mo_a == 0) e[0] = ea_new; else if (mo_a == 1) e[1] = ea_new; else e[2] = ea_new;
            if (mo_b == 0) e[0] = eb_new; else if (mo_b == 1) e[1] = eb_new; else e[2] = eb_new;
            bref(mo.Eoa, oc) = Ea_final;
            bref(mo.Eob, oc) = Eb_final;
            const C Ea_mid = half * (Ea_drive + Ea_final);
            const C Eb_mid = half * (Eb_drive + Eb_final);
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const C C1 = mo_poles[3 * p], C2 = mo_poles[3 * p + 1], C3 = mo_poles[3 * p + 2];
                const C beta = C3 * f * qeff;
                const C Pa_new = fmaT<C>(C1, m.Pa[p], fmaT<C>(C2, m.Ja[p], beta * Eb_old)) + beta * Eb_mid;
                bref(mo.Pa, oc + p * stC) = Pa_new;
                bref(mo.Ja, oc + p * stC) = (Pa_new - m.Pa[p]) * mo_inv_alpha - m.Ja[p];
                const C Pb_new = fmaT<C>(C1, m.Pb[p], fmaT<C>(C2, m.Jb[p], beta * Ea_old)) + beta * Ea_mid;
                bref(mo.Pb, oc + p * stC) = Pb_new;
                bref(mo.Jb, oc + p * stC) = (Pb_new - m.Pb[p]) * mo_inv_alpha - m.Jb[p];
            }
        }
    }
    bref(Ex, oe) = e[0];
    bref(Ey, oe) = e[1];
    bref(Ez, oe) = e[2];
    // pabs_kernel on film cell n
    const double ex = (double)e[0], ey = (double)e[1], ez = (double)e[2];
    const double pabs = ADE_WORK ? jl_max(ex * js[0] + ey * js[1] + ez * js[2], 0.0)
                                 : pabs_factor * (ex * ex + ey * ey + ez * ez);
    const double pdt = pabs * dt;
    bref(U_abs, od) = U_old + pdt;
    bref(window, od) = W_old + pdt;
    if (pabs > peak_old) bref(peak, od) = pabs;
}

template <typename S, typename C, typename T>
static int launch_epilogue(const mp_yee *g, mp_ade *const ade[3], const mp_mo *mo, int mo_a, int mo_b,
                           const void *iea, const void *ieb, mp_film *f, int has_ade_work, double pabs_factor,
                           double Q_TM, double Q_RE, double dt) {
    const int64_t N = f->n;
    AdeView<C, T> v[3];
    for (int c = 0; c < 3; ++c) {
        // the scales of launch_ade (P, J) and launch_pabs (J), formed the same way
        const double sc = state_scale(ade[c]);
        const double jsc = reduced(ade[c]->state_dtype) ? sc * dt / 2.0 : 1.0;
        v[c] = AdeView<C, T>{(T *)ade[c]->P, (T *)ade[c]->J, (C *)ade[c]->E_old, (const C *)ade[c]->fill,
                             (const C *)ade[c]->inv_eps, (float)sc, (float)(1.0 / sc), (float)jsc,
                             (float)(1.0 / jsc)};
    }
    MoView<C> m{};
    if (mo)
        m = MoView<C>{(C *)mo->Pa_from_b, (C *)mo->Ja_from_b, (C *)mo->Pb_from_a, (C *)mo->Jb_from_a,
                      (C *)mo->E_old_a, (C *)mo->E_old_b, (const C *)mo->fill, (const C *)f->m_TM_x,
                      (const C *)f->m_RE_x};
    // the separate launchers form these as CT(2)/CT(dt) (ADE) and 2/CT(dt) (MO): the same value
    const C inv_alpha = (C)2 / (C)dt;
    const C mo_inv_alpha = sizeof(C) == 4 ? (C)(2.0f / (float)dt) : (C)(2.0 / dt);
    KernelTimer t("film_epilogue");
#define EP_ARGS (S *)g->ex, (S *)g->ey, (S *)g->ez, f->cells, v[0], v[1], v[2], (const C *)ade[0]->poles,       \
                ade[0]->npoles, inv_alpha, m, mo ? (const C *)mo->poles : nullptr, mo ? mo->npoles : 0, mo_a, mo_b, \
                (const S *)iea, (const S *)ieb, (C)Q_TM, (C)Q_RE, mo_inv_alpha, f->U_abs, f->window_sum, f->peak, \
                pabs_factor, dt, N
    const unsigned grid = cdiv(N, 256);
    // the two-pole kernel addresses every array by a 32-bit byte offset: the E volumes and
    // two pole blocks of the widest per-cell array (8-byte U_abs / window / peak) must fit
    const int64_t cells_total = g->nx * g->ny * g->nz;
    const bool offsets_fit = cells_total * (int64_t)sizeof(S) < ((int64_t)1 << 32) &&
                             2 * N * (int64_t)std::max<size_t>(sizeof(C), 8) < ((int64_t)1 << 32);
    if (ade[0]->npoles == 2 && (!mo || mo->npoles == 2) && offsets_fit) {
        if (mo) {
            if (has_ade_work) film_epilogue_np_kernel<S, C, T, true, true, 2><<<grid, 256, 0, stream()>>>(EP_ARGS);
            else film_epilogue_np_kernel<S, C, T, true, false, 2><<<grid, 256, 0, stream()>>>(EP_ARGS);
        } else {
            if (has_ade_work) film_epilogue_np_kernel<S, C, T, false, true, 2><<<grid, 256, 0, stream()>>>(EP_ARGS);
            else film_epilogue_np_kernel<S, C, T, false, false, 2><<<grid, 256, 0, stream()>>>(EP_ARGS);
        }
    } else if (mo) {
        if (has_ade_work) film_epilogue_kernel<S, C, T, true, true><<<grid, 256, 0, stream()>>>(EP_ARGS);
        else film_epilogue_kernel<S, C, T, true, false><<<grid, 256, 0, stream()>>>(EP_ARGS);
    } else {
        if (has_ade_work) film_epilogue_kernel<S, C, T, false, true><<<grid, 256, 0, stream()>>>(EP_ARGS);
        else film_epilogue_kernel<S, C, T, false, false><<<grid, 256, 0, stream()>>>(EP_ARGS);
    }
#undef EP_ARGS
    MP_LAUNCH_CHECK("film_epilogue_kernel");
    return MP_OK;
}

}  // namespace mp

using namespace mp;

MP_API int mp_patch_E_dispersive(void *E, mp_ade *s, int32_t field_dtype, double dt) {
    if (!s) return fail_arg("mp_patch_E_dispersive: null state");
    if (s->n <= 0 || s->npoles <= 0) return MP_OK;  // Kernels.jl:599
    if (s->npoles > MAXP) return fail_arg("mp_patch_E_dispersive: at most %d poles", MAXP);
    if (!E || !s->P || !s->J || !s->E_old || !s->idx || !s->fill || !s->inv_eps || !s->poles)
        return fail_arg("mp_patch_E_dispersive: null array");
    const int32_t sd = s->state_dtype;
    if (field_dtype == MP_F32 && s->dtype == MP_F32) {
        if (sd == MP_F32) return launch_ade<float, float, float>(E, s, dt);
        if (sd == MP_BF16) return launch_ade<float, float, bf16>(E, s, dt);
        if (sd == MP_F16) return launch_ade<float, float, f16s>(E, s, dt);
    } else if (field_dtype == MP_F64 && s->dtype == MP_F64 && sd == MP_F64) {
        return launch_ade<double, double, double>(E, s, dt);
    } else if (field_dtype == MP_F32 && s->dtype == MP_F64 && sd == MP_F64) {
        return launch_ade<float, double, double>(E, s, dt);
    }
    return fail_arg("mp_patch_E_dispersive: unsupported dtypes field=%d ct=%d state=%d", field_dtype,
                    s->dtype, sd);
}

MP_API int mp_patch_E_mo_gyration(void *Ea, void *Eb, mp_mo *s, const void *iea, const void *ieb,
                                  int32_t field_dtype, const void *m_TM, const void *m_RE,
                                  double Q_TM, double Q_RE, double dt) {
    if (!s) return fail_arg("mp_patch_E_mo_gyration: null state");
    if (s->n <= 0 || s->npoles <= 0) return MP_OK;  // Kernels.jl:612
    if (!Ea || !Eb || !iea || !ieb || !m_TM || !m_RE || !s->idx || !s->mpos || !s->fill || !s->poles)
        return fail_arg("mp_patch_E_mo_gyration: null array");
    const int64_t N = s->n;
    KernelTimer t("mo_gyration");
    if (field_dtype == MP_F32 && s->dtype == MP_F32) {
        mo_gyration_kernel<float, float><<<cdiv(N, 256), 256, 0, stream()>>>(
            (float *)Ea, (float *)Eb, (float *)s->Pa_from_b, (float *)s->Ja_from_b,
            (float *)s->Pb_from_a, (float *)s->Jb_from_a, (float *)s->E_old_a, (float *)s->E_old_b,
            s->idx, (const float *)s->fill, s->mpos, (const float *)iea, (const float *)ieb,
            (const float *)m_TM, (const float *)m_RE, (const float *)s->poles, s->npoles, N,
            (float)Q_TM, (float)Q_RE, 2.0f / (float)dt);
    } else if (field_dtype == MP_F64 && s->dtype == MP_F64) {
        mo_gyration_kernel<double, double><<<cdiv(N, 256), 256, 0, stream()>>>(
            (double *)Ea, (double *)Eb, (double *)s->Pa_from_b, (double *)s->Ja_from_b,
            (double *)s->Pb_from_a, (double *)s->Jb_from_a, (double *)s->E_old_a, (double *)s->E_old_b,
            s->idx, (const double *)s->fill, s->mpos, (const double *)iea, (const double *)ieb,
            (const double *)m_TM, (const double *)m_RE, (const double *)s->poles, s->npoles, N, Q_TM,
            Q_RE, 2.0 / dt);
    } else if (field_dtype == MP_F32 && s->dtype == MP_F64) {
        mo_gyration_kernel<float, double><<<cdiv(N, 256), 256, 0, stream()>>>(
            (float *)Ea, (float *)Eb, (double *)s->Pa_from_b, (double *)s->Ja_from_b,
            (double *)s->Pb_from_a, (double *)s->Jb_from_a, (double *)s->E_old_a, (double *)s->E_old_b,
            s->idx, (const double *)s->fill, s->mpos, (const float *)iea, (const float *)ieb,
            (const double *)m_TM, (const double *)m_RE, (const double *)s->poles, s->npoles, N, Q_TM,
            Q_RE, 2.0 / dt);
    } else {
        return fail_arg("mp_patch_E_mo_gyration: unsupported dtypes");
    }
    MP_LAUNCH_CHECK("mo_gyration_kernel");
    return MP_OK;
}

MP_API int mp_pabs_accumulate(mp_film *f, const mp_yee *g, const mp_ade *ax, const mp_ade *ay,
                              const mp_ade *az, int32_t has_ade_work, double pabs_factor, double dt) {
    if (!f || !g) return fail_arg("mp_pabs_accumulate: null argument");
    if (f->n <= 0) return MP_OK;  // Kernels.jl:629
    if (has_ade_work) {
        if (!ax || !ay || !az || ax->n <= 0) has_ade_work = 0;  // Kernels.jl:632-634
        else if (ax->npoles != ay->npoles || ax->npoles != az->npoles)
            return fail_arg("mp_pabs_accumulate: ADE lists disagree on npoles");
    }
    if (has_ade_work && (!f->pos_x || !f->pos_y || !f->pos_z))
        return fail_arg("mp_pabs_accumulate: ade_work needs pos_x/pos_y/pos_z");
    const int32_t fd = g->field_dtype;
    if (!has_ade_work) {
        if (fd == MP_F32) return launch_pabs<float, float, float>(f, g, ax, ay, az, 0, pabs_factor, dt);
        if (fd == MP_F64) return launch_pabs<double, double, double>(f, g, ax, ay, az, 0, pabs_factor, dt);
        return fail_arg("mp_pabs_accumulate: unsupported field dtype");
    }
    const int32_t ct = ax->dtype, sd = ax->state_dtype;
    if (fd == MP_F32 && ct == MP_F32) {
        if (sd == MP_F32) return launch_pabs<float, float, float>(f, g, ax, ay, az, 1, pabs_factor, dt);
        if (sd == MP_BF16) return launch_pabs<float, float, bf16>(f, g, ax, ay, az, 1, pabs_factor, dt);
        if (sd == MP_F16) return launch_pabs<float, float, f16s>(f, g, ax, ay, az, 1, pabs_factor, dt);
    } else if (fd == MP_F64 && ct == MP_F64 && sd == MP_F64) {
        return launch_pabs<double, double, double>(f, g, ax, ay, az, 1, pabs_factor, dt);
    } else if (fd == MP_F32 && ct == MP_F64 && sd == MP_F64) {
        return launch_pabs<float, double, double>(f, g, ax, ay, az, 1, pabs_factor, dt);
    }
    return fail_arg("mp_pabs_accumulate: unsupported dtypes");
}

static int component_of(const mp_yee *g, const void *E) {
    return E == g->ex ? 0 : (E == g->ey ? 1 : (E == g->ez ? 2 : -1));
}

MP_API int mp_film_epilogue(const mp_yee *g, mp_ade *ax, mp_ade *ay, mp_ade *az, mp_mo *mo, void *mo_ea,
                            void *mo_eb, const void *mo_iea, const void *mo_ieb, mp_film *f, int32_t has_ade_work,
                            double pabs_factor, double Q_TM, double Q_RE, double dt) {
    if (!g || !ax || !ay || !az || !f) return fail_arg("mp_film_epilogue: null argument");
    if (f->n <= 0) return MP_OK;
    mp_ade *ade[3] = {ax, ay, az};
    const int64_t N = f->n;
    for (int c = 0; c < 3; ++c) {
        const mp_ade *a = ade[c];
        if (a->n != N) return fail_arg("mp_film_epilogue: ADE list %d has %lld entries, the film %lld", c,
                                       (long long)a->n, (long long)N);
        if (a->npoles != ax->npoles) return fail_arg("mp_film_epilogue: ADE lists disagree on npoles");
        if (a->npoles < 1 || a->npoles > MAXP) return fail_arg("mp_film_epilogue: 1..%d poles", MAXP);
        if (a->dtype != ax->dtype || a->state_dtype != ax->state_dtype || state_scale(a) != state_scale(ax))
            return fail_arg("mp_film_epilogue: the three ADE lists must share compute type, storage and scale");
        if (a->state_dtype != a->dtype && !(a->dtype == MP_F32 && reduced(a->state_dtype)))
            return fail_arg("mp_film_epilogue: ADE state stored in the compute type, or BF16/F16 with FP32");
        if (!a->P || !a->J || !a->E_old || !a->fill || !a->inv_eps || !a->poles)
            return fail_arg("mp_film_epilogue: null ADE array");
    }
    if (!f->cells || !f->U_abs || !f->window_sum || !f->peak) return fail_arg("mp_film_epilogue: null film array");
    int mo_a = -1, mo_b = -1;
    if (mo && mo->n > 0 && mo->npoles > 0) {
        if (mo->n != N) return fail_arg("mp_film_epilogue: MO list must be the film list");
        if (mo->dtype != ax->dtype) return fail_arg("mp_film_epilogue: MO dtype must equal the ADE compute type");
        mo_a = component_of(g, mo_ea);
        mo_b = component_of(g, mo_eb);
        if (mo_a < 0 || mo_b < 0 || mo_a == mo_b)
            return fail_arg("mp_film_epilogue: the MO pair must be two distinct E volumes of the state");
        if (!mo_iea || !mo_ieb || !mo->Pa_from_b || !mo->Ja_from_b || !mo->Pb_from_a || !mo->Jb_from_a ||
            !mo->E_old_a || !mo->E_old_b || !mo->fill || !mo->poles || !f->m_TM_x || !f->m_RE_x)
            return fail_arg("mp_film_epilogue: null MO array");
    } else {
        mo = nullptr;
    }
    const int32_t fd = g->field_dtype, ct = ax->dtype, sd = ax->state_dtype;
#define EP_CALL(S, C, T) \
    launch_epilogue<S, C, T>(g, ade, mo, mo_a, mo_b, mo_iea, mo_ieb, f, has_ade_work, pabs_factor, Q_TM, Q_RE, dt)
    if (fd == MP_F32 && ct == MP_F32) {
        if (sd == MP_BF16) return EP_CALL(float, float, bf16);
        if (sd == MP_F16) return EP_CALL(float, float, f16s);
        return EP_CALL(float, float, float);
    }
    if (fd == MP_F64 && ct == MP_F64) return EP_CALL(double, double, double);
    if (fd == MP_F32 && ct == MP_F64) return EP_CALL(float, double, double);
#undef EP_CALL
    return fail_arg("mp_film_epilogue: unsupported dtypes field=%d ct=%d", fd, ct);
}
