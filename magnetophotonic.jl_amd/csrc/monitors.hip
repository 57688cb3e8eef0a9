// Monitor kernels: running FP64 DFT + net-flux integral on monitor planes, mode-weighted
// plane traces, plane Poynting flux, film reductions and the non-finite scan.
//
// Reference: probe_phase_factors_kernel! / probe_dft_accumulate_kernel! /
// probe_trace_plane_kernel! (src/fdtd/Kernels.jl:420-478) with _record_probe_device!
// (src/sim/ProbeReadout.jl:274-301); record!(::FluxMonitor) (src/sim/Monitors.jl:263-341);
// FilmAverage / SwitchedFraction (Monitors.jl:487-578); NaNGuard (Monitors.jl:580-606).
// Reductions are two-pass (fixed per-block partials, then one ordered final sum) so
// they are deterministic run to run.
#include <algorithm>
#include <mutex>

#include "common.h"
#include "monitors.h"

namespace mp {

struct PlaneDesc {
    int32_t axis;
    int64_t index, n1, n2;
    int64_t off;  // first cell of this plane in a snapshot slot
    double *ea, *eb, *ha, *hb, *energy;
};
struct PlaneSet {
    PlaneDesc p[MP_MAX_PLANES];
    int32_t count;
};

__device__ __forceinline__ int64_t plane_lin(int axis, int64_t idx, int64_t a, int64_t b, int64_t nx, int64_t ny) {
    if (axis == 0) return idx + nx * (a + ny * b);   // (j, k)
    if (axis == 1) return b + nx * (idx + ny * a);   // (k, i)
    return a + nx * (b + ny * idx);                  // (i, j)
}

__global__ void phase_factors_kernel(double *phase, const double *omega, int nb, double t, double dt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    const double ph = omega[q] * t;
    phase[2 * q] = cos(ph) * dt;
    phase[2 * q + 1] = -sin(ph) * dt;
}

template <typename S>
__global__ __launch_bounds__(256) void dft_accumulate_kernel(
    PlaneSet ps, const S *Ex, const S *Ey, const S *Ez, const S *Hx, const S *Hy, const S *Hz,
    const double *__restrict__ phase, int nb, double dt, int64_t nx, int64_t ny) {
    const PlaneDesc &pl = ps.p[blockIdx.y];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = pl.n1 * pl.n2;
    if (c >= np) return;
    const int64_t a = c % pl.n1, b = c / pl.n1;
    const int64_t o = plane_lin(pl.axis, pl.index, a, b, nx, ny);
    const S *Ea, *Eb, *Ha, *Hb;
    if (pl.axis == 0) { Ea = Ey; Eb = Ez; Ha = Hy; Hb = Hz; }
    else if (pl.axis == 1) { Ea = Ez; Eb = Ex; Ha = Hz; Hb = Hx; }
    else { Ea = Ex; Eb = Ey; Ha = Hx; Hb = Hy; }
    const double ea = (double)Ea[o], eb = (double)Eb[o], ha = (double)Ha[o], hb = (double)Hb[o];
    pl.energy[c] += (ea * hb - eb * ha) * dt;
    for (int q = 0; q < nb; ++q) {
        const double pr = phase[2 * q], pi = phase[2 * q + 1];
        const int64_t z = 2 * (c + np * q);
        double2 *ea2 = (double2 *)(pl.ea + z), *eb2 = (double2 *)(pl.eb + z);
        double2 *ha2 = (double2 *)(pl.ha + z), *hb2 = (double2 *)(pl.hb + z);
        double2 v;
        v = *ea2; v.x += ea * pr; v.y += ea * pi; *ea2 = v;
        v = *eb2; v.x += eb * pr; v.y += eb * pi; *eb2 = v;
        v = *ha2; v.x += ha * pr; v.y += ha * pi; *ha2 = v;
        v = *hb2; v.x += hb * pr; v.y += hb * pi; *hb2 = v;
    }
}

// ---- deterministic reductions -------------------------------------------------------------
// ---- batched DFT record (used by the step driver) ------------------------------------
// Each step stores the 4 tangential fields of every plane cell in one slot of a snapshot
// buffer and the step's phase factors in one slot of a phase buffer; every DFT_BATCH
// steps (and when the driver returns) one pass adds the stored steps to the FP64
// accumulators in step order.  The additions are the per-step ones in the same order, so
// the accumulators are bit-identical to recording every step, with the accumulators read
// and written once per batch instead of once per step.
constexpr int DFT_MAXNB = 8;         // bins the batched record supports
constexpr int DFT_MAXNB_STEPS = 64;  // longest batch

template <typename S> struct Quad { S a, b, c, d; };

// The step's snapshot of every plane's tangential fields; block (0, 0) also writes the
// step's phase slot, the nb factors (cos wt * dt, -sin wt * dt).
template <typename S>
__global__ __launch_bounds__(256) void dft_snapshot_kernel(PlaneSet ps, const S *Ex, const S *Ey, const S *Ez,
                                                           const S *Hx, const S *Hy, const S *Hz,
                                                           Quad<S> *__restrict__ snap, int64_t nx, int64_t ny,
                                                           double *__restrict__ phase,
                                                           const double *__restrict__ omega, int nb, double t,
                                                           double dt) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && (int)threadIdx.x < nb) {
        const int q = threadIdx.x;
        const double ph = omega[q] * t;
        phase[2 * q] = cos(ph) * dt;
        phase[2 * q + 1] = -sin(ph) * dt;
    }
    const PlaneDesc &pl = ps.p[blockIdx.y];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = pl.n1 * pl.n2;
    if (c >= np) return;
    const int64_t a = c % pl.n1, b = c / pl.n1;
    const int64_t o = plane_lin(pl.axis, pl.index, a, b, nx, ny);
    const S *Ea, *Eb, *Ha, *Hb;
    if (pl.axis == 0) { Ea = Ey; Eb = Ez; Ha = Hy; Hb = Hz; }
    else if (pl.axis == 1) { Ea = Ez; Eb = Ex; Ha = Hz; Hb = Hx; }
    else { Ea = Ex; Eb = Ey; Ha = Hx; Hb = Hy; }
    snap[pl.off + c] = Quad<S>{Ea[o], Eb[o], Ha[o], Hb[o]};
}


template <typename S, int NB>
__global__ __launch_bounds__(256) void dft_batch_kernel(PlaneSet ps, const Quad<S> *__restrict__ snap,
                                                        int64_t slot_stride, const double *__restrict__ phase,
                                                        int nsteps, double dt) {
    // the batch's phase factors, read by every cell, staged once per workgroup
    __shared__ double ph[DFT_MAXNB_STEPS][NB][2];
    for (int i = threadIdx.x; i < nsteps * NB * 2; i += blockDim.x) {
        const int st = i / (NB * 2), r = i % (NB * 2);
        ph[st][r / 2][r % 2] = phase[2 * DFT_MAXNB * st + r];
    }
    __syncthreads();
    const PlaneDesc &pl = ps.p[blockIdx.y];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = pl.n1 * pl.n2;
    if (c >= np) return;
    const Quad<S> *sp = snap + pl.off + c;
    double acc[4][NB][2];
    double en = pl.energy[c];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int64_t z = 2 * (c + np * q);
        const double2 va = *(const double2 *)(pl.ea + z), vb = *(const double2 *)(pl.eb + z);
        const double2 ha = *(const double2 *)(pl.ha + z), hb = *(const double2 *)(pl.hb + z);
        acc[0][q][0] = va.x; acc[0][q][1] = va.y;
        acc[1][q][0] = vb.x; acc[1][q][1] = vb.y;
        acc[2][q][0] = ha.x; acc[2][q][1] = ha.y;
        acc[3][q][0] = hb.x; acc[3][q][1] = hb.y;
    }
    for (int s0 = 0; s0 < nsteps; s0 += 4) {
        Quad<S> f[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (s0 + u < nsteps) f[u] = sp[(s0 + u) * slot_stride];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (s0 + u >= nsteps) break;
            const double ea = (double)f[u].a, eb = (double)f[u].b, ha = (double)f[u].c, hb = (double)f[u].d;
            en += (ea * hb - eb * ha) * dt;
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const double pr = ph[s0 + u][q][0], pi = ph[s0 + u][q][1];
                acc[0][q][0] += ea * pr; acc[0][q][1] += ea * pi;
                acc[1][q][0] += eb * pr; acc[1][q][1] += eb * pi;
                acc[2][q][0] += ha * pr; acc[2][q][1] += ha * pi;
                acc[3][q][0] += hb * pr; acc[3][q][1] += hb * pi;
            }
        }
    }
    pl.energy[c] = en;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int64_t z = 2 * (c + np * q);
        *(double2 *)(pl.ea + z) = make_double2(acc[0][q][0], acc[0][q][1]);
        *(double2 *)(pl.eb + z) = make_double2(acc[1][q][0], acc[1][q][1]);
        *(double2 *)(pl.ha + z) = make_double2(acc[2][q][0], acc[2][q][1]);
        *(double2 *)(pl.hb + z) = make_double2(acc[3][q][0], acc[3][q][1]);
    }
}

template <typename S>
static void launch_dft_batch(dim3 grid, const PlaneSet &ps, const void *snap, int64_t total, const double *phase,
                             int n, int nb, double dt) {
    const Quad<S> *q = (const Quad<S> *)snap;
    switch (nb) {
    case 1: dft_batch_kernel<S, 1><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    case 2: dft_batch_kernel<S, 2><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    case 3: dft_batch_kernel<S, 3><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    case 4: dft_batch_kernel<S, 4><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    case 5: dft_batch_kernel<S, 5><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    case 6: dft_batch_kernel<S, 6><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    case 7: dft_batch_kernel<S, 7><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    default: dft_batch_kernel<S, 8><<<grid, 256, 0, stream()>>>(ps, q, total, phase, n, dt); break;
    }
}

constexpr int RED_BLOCKS = 256;
constexpr int RED_THREADS = 256;
constexpr int RED_MAXV = 16;

static double *g_partials = nullptr;  // RED_BLOCKS * RED_MAXV doubles
static std::mutex g_red_mu;

static int partials(double **out) {
    std::lock_guard<std::mutex> lk(g_red_mu);
    if (!g_partials) MP_HIP(hipMalloc((void **)&g_partials, sizeof(double) * RED_BLOCKS * RED_MAXV));
    *out = g_partials;
    return MP_OK;
}

template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double *dst /* NV */) {
    __shared__ double sh[RED_THREADS];
    for (int q = 0; q < NV; ++q) {
        sh[threadIdx.x] = v[q];
        __syncthreads();
        for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) dst[q] = sh[0];
        __syncthreads();
    }
}

template <typename S>
__global__ __launch_bounds__(RED_THREADS) void trace_partial_kernel(
    const S *Ea, const S *Eb, int axis, int64_t ipre, int64_t ipost, const double *mode_w,
    const double *area, int64_t n1, int64_t n2, int64_t nx, int64_t ny, double *part) {
    double v[4] = {0, 0, 0, 0};
    const int64_t np = n1 * n2;
    for (int64_t c = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; c < np; c += (int64_t)RED_BLOCKS * RED_THREADS) {
        const int64_t a = c % n1, b = c / n1;
        const double w = mode_w[c] * area[c];
        const int64_t o1 = plane_lin(axis, ipre, a, b, nx, ny), o2 = plane_lin(axis, ipost, a, b, nx, ny);
        v[0] += (double)Ea[o1] * w;
        v[1] += (double)Eb[o1] * w;
        v[2] += (double)Ea[o2] * w;
        v[3] += (double)Eb[o2] * w;
    }
    block_sum_store<4>(v, part + 4 * blockIdx.x);
}

template <typename S>
__global__ __launch_bounds__(RED_THREADS) void flux_partial_kernel(
    const S *Ea, const S *Eb, const S *Ha, const S *Hb, int axis, int64_t idx, const double *area,
    int64_t n1, int64_t n2, int64_t nx, int64_t ny, double *part) {
    double v[1] = {0};
    const int64_t np = n1 * n2;
    for (int64_t c = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; c < np; c += (int64_t)RED_BLOCKS * RED_THREADS) {
        const int64_t a = c % n1, b = c / n1;
        const int64_t o = plane_lin(axis, idx, a, b, nx, ny);
        v[0] += ((double)Ea[o] * (double)Hb[o] - (double)Eb[o] * (double)Ha[o]) * area[c];
    }
    block_sum_store<1>(v, part + blockIdx.x);
}

template <typename L>
__global__ __launch_bounds__(RED_THREADS) void film_partial_kernel(
    const L *mTx, const L *mTy, const L *mTz, const L *mRx, const L *mRy, const L *mRz,
    const L *Te, const L *Tl, int64_t N, double *part) {
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, -INFINITY};
    for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < N; n += (int64_t)RED_BLOCKS * RED_THREADS) {
        v[0] += (double)mTx[n]; v[1] += (double)mTy[n]; v[2] += (double)mTz[n];
        v[3] += (double)mRx[n]; v[4] += (double)mRy[n]; v[5] += (double)mRz[n];
        const double te = (double)Te[n];
        v[6] += te; v[7] += (double)Tl[n];
        v[8] += (double)mTx[n] < 0.0 ? 1.0 : 0.0;
        v[9] = te > v[9] ? te : v[9];
    }
    double s[9];
    for (int q = 0; q < 9; ++q) s[q] = v[q];
    block_sum_store<9>(s, part + RED_MAXV * blockIdx.x);
    __shared__ double mx[RED_THREADS];
    mx[threadIdx.x] = v[9];
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) mx[threadIdx.x] = mx[threadIdx.x + st] > mx[threadIdx.x] ? mx[threadIdx.x + st] : mx[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[RED_MAXV * blockIdx.x + 9] = mx[0];
}

// Final ordered sum of the partials: out[q] = scale * sum_b part[stride*b + q] (q < nsum),
// out[q] = max_b part[...] for nsum <= q < nv.
__global__ void finalize_kernel(const double *part, int stride, int nsum, int nv, double scale, double *out) {
    const int q = threadIdx.x;
    if (q >= nv) return;
    if (q < nsum) {
        double s = 0.0;
        for (int b = 0; b < RED_BLOCKS; ++b) s += part[stride * b + q];
        out[q] = s * scale;
    } else {
        double m = -INFINITY;
        for (int b = 0; b < RED_BLOCKS; ++b) m = part[stride * b + q] > m ? part[stride * b + q] : m;
        out[q] = m;
    }
}

template <typename T>
__global__ void nonfinite_kernel(const T *x, int64_t n, int32_t *flag) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        bad |= !__builtin_isfinite((double)x[i]);
    if (bad) *flag = 1;
}

static void plane_fields(const mp_yee *g, int axis, const void **Ea, const void **Eb, const void **Ha,
                         const void **Hb, int64_t *n1, int64_t *n2) {
    if (axis == 0) { *Ea = g->ey; *Eb = g->ez; *Ha = g->hy; *Hb = g->hz; *n1 = g->ny; *n2 = g->nz; }
    else if (axis == 1) { *Ea = g->ez; *Eb = g->ex; *Ha = g->hz; *Hb = g->hx; *n1 = g->nz; *n2 = g->nx; }
    else { *Ea = g->ex; *Eb = g->ey; *Ha = g->hx; *Hb = g->hy; *n1 = g->nx; *n2 = g->ny; }
}

// ---- probe shot read-out (ProbeReadout.jl:322-351) ------------------------------------
//
// Grid (SHOT_BLOCKS, nb + 1).  Row q < nb reduces bin q's complex Poynting cross-power
// sum((Ea*conj(Hb) - Eb*conj(Ha)) * dA) on both planes; row nb reduces the net flux
// sum(energy * dA) on both planes and the four mode overlaps sum(dft[., q0] * w * dA) at
// the centre bin (post Ea, post Eb, pre Ea, pre Eb).  Every row reads its own
// contiguous (n1*n2) slab of each DFT array, so loads are coalesced 16-byte pairs.
constexpr int SHOT_BLOCKS = 64;
constexpr int SHOT_NV = 10;

struct ShotPlane {
    const double2 *ea, *eb, *ha, *hb;
    const double *energy;
};

__device__ __forceinline__ void poynting_term(double2 a, double2 b, double2 ha, double2 hb, double dA, double &re,
                                              double &im) {
    // a*conj(hb) - b*conj(ha)
    re += ((a.x * hb.x + a.y * hb.y) - (b.x * ha.x + b.y * ha.y)) * dA;
    im += ((a.y * hb.x - a.x * hb.y) - (b.y * ha.x - b.x * ha.y)) * dA;
}

__global__ __launch_bounds__(RED_THREADS) void shot_partial_kernel(ShotPlane pre, ShotPlane post,
                                                                   const double *mode_w, const double *area,
                                                                   int64_t np, int nb, int q0, double *part) {
    const int q = blockIdx.y;
    double v[SHOT_NV] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t c0 = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x;
    const int64_t cs = (int64_t)SHOT_BLOCKS * RED_THREADS;
    if (q < nb) {
        const int64_t o = (int64_t)q * np;
        for (int64_t c = c0; c < np; c += cs) {
            const double dA = area[c];
            poynting_term(pre.ea[o + c], pre.eb[o + c], pre.ha[o + c], pre.hb[o + c], dA, v[0], v[1]);
            poynting_term(post.ea[o + c], post.eb[o + c], post.ha[o + c], post.hb[o + c], dA, v[2], v[3]);
        }
    } else {
        const int64_t o = (int64_t)q0 * np;
        for (int64_t c = c0; c < np; c += cs) {
            const double dA = area[c];
            v[0] += pre.energy[c] * dA;
            v[1] += post.energy[c] * dA;
            const double w = mode_w[c] * dA;
            const double2 fa = post.ea[o + c], fb = post.eb[o + c], ka = pre.ea[o + c], kb = pre.eb[o + c];
            v[2] += fa.x * w; v[3] += fa.y * w;
            v[4] += fb.x * w; v[5] += fb.y * w;
            v[6] += ka.x * w; v[7] += ka.y * w;
            v[8] += kb.x * w; v[9] += kb.y * w;
        }
    }
    block_sum_store<SHOT_NV>(v, part + SHOT_NV * ((int64_t)q * SHOT_BLOCKS + blockIdx.x));
}

// One thread per output value, partials summed in block order.  Output layout:
// [0, 2nb) Sx_pre (re, im), [2nb, 4nb) Sx_post, then the SHOT_NV values of row nb.
__global__ void shot_finalize_kernel(const double *part, int nb, double *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nout = 4 * nb + SHOT_NV;
    if (t >= nout) return;
    int q, v;
    if (t < 4 * nb) {
        const int plane = t / (2 * nb), r = t % (2 * nb);
        q = r / 2;
        v = 2 * plane + (r & 1);
    } else {
        q = nb;
        v = t - 4 * nb;
    }
    double s = 0.0;
    for (int b = 0; b < SHOT_BLOCKS; ++b) s += part[SHOT_NV * ((int64_t)q * SHOT_BLOCKS + b) + v];
    out[t] = s;
}

// Weighted plane sum of one field component: sum(F[plane cell] * w[cell]), w = 1 when
// weight is null (Monitors.jl:350-368 and :461-473).
template <typename S>
__global__ __launch_bounds__(RED_THREADS) void overlap_partial_kernel(const S *F, int axis, int64_t idx,
                                                                      const double *weight, int64_t n1, int64_t n2,
                                                                      int64_t nx, int64_t ny, double *part) {
    double v[1] = {0};
    const int64_t np = n1 * n2;
    for (int64_t c = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; c < np; c += (int64_t)RED_BLOCKS * RED_THREADS) {
        const int64_t a = c % n1, b = c / n1;
        const double f = (double)F[plane_lin(axis, idx, a, b, nx, ny)];
        v[0] += weight ? f * weight[c] : f;
    }
    block_sum_store<1>(v, part + blockIdx.x);
}


// ---- film averages and switching diagnostics ------------------------------------------------
//
// FilmAverage (Monitors.jl:518-570): per-block sums of the six m components, |m_TM|,
// |m_RE|, the four temperatures and the count of m_TM_x < 0 (13 values).
template <typename L>
__global__ __launch_bounds__(RED_THREADS) void film_average_partial_kernel(
    const L *mTx, const L *mTy, const L *mTz, const L *mRx, const L *mRy, const L *mRz, const L *Te, const L *Tl,
    const L *TsT, const L *TsR, int64_t N, double *part) {
    double v[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < N; n += (int64_t)RED_BLOCKS * RED_THREADS) {
        const double tx = (double)mTx[n], ty = (double)mTy[n], tz = (double)mTz[n];
        const double rx = (double)mRx[n], ry = (double)mRy[n], rz = (double)mRz[n];
        v[0] += tx; v[1] += ty; v[2] += tz;
        v[3] += rx; v[4] += ry; v[5] += rz;
        v[6] += sqrt(tx * tx + ty * ty + tz * tz);
        v[7] += sqrt(rx * rx + ry * ry + rz * rz);
        v[8] += (double)Te[n]; v[9] += (double)Tl[n]; v[10] += (double)TsT[n]; v[11] += (double)TsR[n];
        v[12] += mTx[n] < (L)0 ? 1.0 : 0.0;
    }
    block_sum_store<13>(v, part + RED_MAXV * blockIdx.x);
}

// switching_metrics pass 1 (Monitors.jl:682-727): switched = m_TM_x < 0 and m_RE_x > 0.
// Sums: switched count, TM-reversed count, RE-reversed count, m_TM_x and m_RE_x over
// switched and over unswitched cells (7); then max U_abs and its first index.
template <typename L>
__global__ __launch_bounds__(RED_THREADS) void switch_partial_kernel(const L *mTx, const L *mRx, const double *U,
                                                                      int64_t N, double *part) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    double umax = -INFINITY;
    int64_t imax = N;
    for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < N; n += (int64_t)RED_BLOCKS * RED_THREADS) {
        const double tm = (double)mTx[n], re = (double)mRx[n];
        const bool sw = tm < 0.0 && re > 0.0;
        v[0] += sw ? 1.0 : 0.0;
        v[1] += tm < 0.0 ? 1.0 : 0.0;
        v[2] += re > 0.0 ? 1.0 : 0.0;
        if (sw) { v[3] += tm; v[4] += re; } else { v[5] += tm; v[6] += re; }
        const double u = U[n];
        if (u > umax || imax == N) { umax = u; imax = n; }   // n rises per thread: first max kept
    }
    block_sum_store<7>(v, part + RED_MAXV * blockIdx.x);
    __shared__ double mv[RED_THREADS];
    __shared__ int64_t mi[RED_THREADS];
    mv[threadIdx.x] = umax;
    mi[threadIdx.x] = imax;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const double ov = mv[threadIdx.x + st];
            const int64_t oi = mi[threadIdx.x + st];
            if (oi < N && (mi[threadIdx.x] == N || ov > mv[threadIdx.x] || (ov == mv[threadIdx.x] && oi < mi[threadIdx.x]))) {
                mv[threadIdx.x] = ov;
                mi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[RED_MAXV * blockIdx.x + 7] = mv[0];
        part[RED_MAXV * blockIdx.x + 8] = (double)mi[0];
    }
}

// Ordered sums of the 7 pass-1 values, and the first maximum over blocks: out[0..8].
__global__ void switch_finalize_kernel(const double *part, int64_t N, double *out) {
    const int q = threadIdx.x;
    if (q < 7) {
        double s = 0.0;
        for (int b = 0; b < RED_BLOCKS; ++b) s += part[RED_MAXV * b + q];
        out[q] = s;
    } else if (q == 7) {
        double m = -INFINITY;
        double mi = (double)N;
        for (int b = 0; b < RED_BLOCKS; ++b) {
            const double v = part[RED_MAXV * b + 7], i = part[RED_MAXV * b + 8];
            if (i < (double)N && (mi == (double)N || v > m || (v == m && i < mi))) { m = v; mi = i; }
        }
        out[7] = m;
        out[8] = mi;
    }
}

// switching_metrics pass 2: the core is U_abs >= core_fraction * max(U_abs); counts of core
// and core-and-switched cells, m_TM_x and m_RE_x summed over the core.
template <typename L>
__global__ __launch_bounds__(RED_THREADS) void core_partial_kernel(const L *mTx, const L *mRx, const double *U,
                                                                    int64_t N, const double *stats, double core_fraction,
                                                                    double *part) {
    const double thr = core_fraction * stats[7];
    double v[4] = {0, 0, 0, 0};
    for (int64_t n = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; n < N; n += (int64_t)RED_BLOCKS * RED_THREADS) {
        if (U[n] >= thr) {
            const double tm = (double)mTx[n], re = (double)mRx[n];
            v[0] += 1.0;
            v[1] += (tm < 0.0 && re > 0.0) ? 1.0 : 0.0;
            v[2] += tm;
            v[3] += re;
        }
    }
    block_sum_store<4>(v, part + RED_MAXV * blockIdx.x);
}

}  // namespace mp

using namespace mp;

namespace mp {

// The plane set of one record, validated; off = running offset of each plane in a
// snapshot slot, *total = cells of all planes.
static int build_planes(const mp_yee *g, const mp_plane *planes, int32_t nplanes, PlaneSet &ps, int64_t *maxnp,
                        int64_t *total, const char *who) {
    if (nplanes > MP_MAX_PLANES) return fail_arg("%s: at most %d planes", who, MP_MAX_PLANES);
    ps = PlaneSet{};
    ps.count = nplanes;
    *maxnp = 0;
    *total = 0;
    for (int p = 0; p < nplanes; ++p) {
        const mp_plane &src = planes[p];
        if (src.axis < 0 || src.axis > 2) return fail_arg("%s: bad plane axis", who);
        const int64_t nax = src.axis == 0 ? g->nx : (src.axis == 1 ? g->ny : g->nz);
        if (src.index < 0 || src.index >= nax) return fail_arg("%s: plane index out of range", who);
        if (!src.dft_ea || !src.dft_eb || !src.dft_ha || !src.dft_hb || !src.energy)
            return fail_arg("%s: null plane buffer", who);
        PlaneDesc &d = ps.p[p];
        d.axis = src.axis;
        d.index = src.index;
        d.n1 = src.axis == 0 ? g->ny : (src.axis == 1 ? g->nz : g->nx);
        d.n2 = src.axis == 0 ? g->nz : (src.axis == 1 ? g->nx : g->ny);
        d.off = *total;
        d.ea = src.dft_ea; d.eb = src.dft_eb; d.ha = src.dft_ha; d.hb = src.dft_hb; d.energy = src.energy;
        *maxnp = std::max(*maxnp, d.n1 * d.n2);
        *total += d.n1 * d.n2;
    }
    return MP_OK;
}

struct DftBatch {
    void *snap = nullptr;
    size_t snap_bytes = 0;
    double *phase = nullptr;
    int count = 0;
};
static DftBatch g_dft;

// Steps per batch: 32 (batches of 64 steps measured no faster and double the buffer).
int dft_batch_size() { return 32; }

int dft_flush(const mp_yee *g, const mp_plane *planes, int32_t nplanes, int32_t nb, double dt) {
    if (g_dft.count == 0) return MP_OK;
    PlaneSet ps;
    int64_t maxnp, total;
    if (int e = build_planes(g, planes, nplanes, ps, &maxnp, &total, "dft flush")) return e;
    const int n = g_dft.count;
    g_dft.count = 0;
    dim3 grid(cdiv(maxnp, 256), nplanes);
    KernelTimer t("dft_batch");
    if (g->field_dtype == MP_F32) launch_dft_batch<float>(grid, ps, g_dft.snap, total, g_dft.phase, n, nb, dt);
    else launch_dft_batch<double>(grid, ps, g_dft.snap, total, g_dft.phase, n, nb, dt);
    MP_LAUNCH_CHECK("dft_batch_kernel");
    return MP_OK;
}

bool dft_batch_full() { return g_dft.count == dft_batch_size(); }

int dft_record_batched(const mp_yee *g, const mp_plane *planes, int32_t nplanes, const double *omega_bins,
                       double *phase, int32_t nb, double t, double dt) {
    const int K = dft_batch_size();
    if (K <= 1 || nb > DFT_MAXNB || (g->field_dtype != MP_F32 && g->field_dtype != MP_F64))
        return mp_probe_dft_accumulate(g, planes, nplanes, omega_bins, phase, nb, t, dt);
    if (!g || (!planes && nplanes > 0)) return fail_arg("dft record: null argument");
    if (nplanes <= 0 || nb <= 0) return MP_OK;
    if (!omega_bins) return fail_arg("dft record: null bins");
    // a full batch the caller has not flushed yet goes first
    if (g_dft.count == K)
        if (int e = dft_flush(g, planes, nplanes, nb, dt)) return e;
    PlaneSet ps;
    int64_t maxnp, total;
    if (int e = build_planes(g, planes, nplanes, ps, &maxnp, &total, "dft record")) return e;
    const size_t qsz = g->field_dtype == MP_F32 ? sizeof(Quad<float>) : sizeof(Quad<double>);
    const size_t need = qsz * (size_t)total * (size_t)K;
    if (g_dft.snap_bytes < need) {
        if (g_dft.count) return fail_arg("dft record: plane set changed inside a batch");
        if (g_dft.snap) MP_HIP(hipFree(g_dft.snap));
        g_dft.snap = nullptr;
        g_dft.snap_bytes = 0;
        MP_HIP(hipMalloc(&g_dft.snap, need));
        g_dft.snap_bytes = need;
    }
    if (K > DFT_MAXNB_STEPS) return fail_arg("dft record: batch longer than %d steps", DFT_MAXNB_STEPS);
    if (!g_dft.phase) MP_HIP(hipMalloc((void **)&g_dft.phase, sizeof(double) * 2 * DFT_MAXNB * DFT_MAXNB_STEPS));
    {
        dim3 grid(cdiv(maxnp, 256), nplanes);
        double *slot = g_dft.phase + 2 * DFT_MAXNB * g_dft.count;
        KernelTimer t1("dft_snapshot");
        if (g->field_dtype == MP_F32)
            dft_snapshot_kernel<float><<<grid, 256, 0, stream()>>>(
                ps, (const float *)g->ex, (const float *)g->ey, (const float *)g->ez, (const float *)g->hx,
                (const float *)g->hy, (const float *)g->hz, (Quad<float> *)g_dft.snap + total * g_dft.count, g->nx,
                g->ny, slot, omega_bins, nb, t, dt);
        else
            dft_snapshot_kernel<double><<<grid, 256, 0, stream()>>>(
                ps, (const double *)g->ex, (const double *)g->ey, (const double *)g->ez, (const double *)g->hx,
                (const double *)g->hy, (const double *)g->hz, (Quad<double> *)g_dft.snap + total * g_dft.count,
                g->nx, g->ny, slot, omega_bins, nb, t, dt);
        MP_LAUNCH_CHECK("dft_snapshot_kernel");
    }
    ++g_dft.count;   // a full batch waits for dft_flush (the step driver runs it where it is cheapest)
    return MP_OK;
}

}  // namespace mp

MP_API int mp_probe_dft_accumulate(const mp_yee *g, const mp_plane *planes, int32_t nplanes,
                                   const double *omega_bins, double *phase, int32_t nb, double t,
                                   double dt) {
    if (!g || (!planes && nplanes > 0)) return fail_arg("mp_probe_dft_accumulate: null argument");
    if (nplanes <= 0 || nb <= 0) return MP_OK;  // Kernels.jl:677-679
    if (!omega_bins || !phase) return fail_arg("mp_probe_dft_accumulate: null bins/phase buffer");
    PlaneSet ps;
    int64_t maxnp, total;
    if (int e = build_planes(g, planes, nplanes, ps, &maxnp, &total, "mp_probe_dft_accumulate")) return e;
    {
        KernelTimer t0("dft_phase");
        phase_factors_kernel<<<cdiv(nb, 64), 64, 0, stream()>>>(phase, omega_bins, nb, t, dt);
        MP_LAUNCH_CHECK("phase_factors_kernel");
    }
    dim3 grid(cdiv(maxnp, 256), nplanes);
    KernelTimer t1("dft_accumulate");
    if (g->field_dtype == MP_F32) {
        dft_accumulate_kernel<float><<<grid, 256, 0, stream()>>>(ps, (const float *)g->ex, (const float *)g->ey,
            (const float *)g->ez, (const float *)g->hx, (const float *)g->hy, (const float *)g->hz, phase, nb,
            dt, g->nx, g->ny);
    } else if (g->field_dtype == MP_F64) {
        dft_accumulate_kernel<double><<<grid, 256, 0, stream()>>>(ps, (const double *)g->ex, (const double *)g->ey,
            (const double *)g->ez, (const double *)g->hx, (const double *)g->hy, (const double *)g->hz, phase, nb,
            dt, g->nx, g->ny);
    } else {
        return fail_arg("mp_probe_dft_accumulate: unsupported field dtype");
    }
    MP_LAUNCH_CHECK("dft_accumulate_kernel");
    return MP_OK;
}

MP_API int mp_probe_trace_plane(const mp_yee *g, int32_t axis, int64_t i_pre, int64_t i_post,
                                const double *mode_w, const double *area, double *out) {
    if (!g || !mode_w || !area || !out) return fail_arg("mp_probe_trace_plane: null argument");
    if (axis < 0 || axis > 2) return fail_arg("mp_probe_trace_plane: bad axis");
    const int64_t nax = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    if (i_pre < 0 || i_pre >= nax || i_post < 0 || i_post >= nax) return fail_arg("mp_probe_trace_plane: plane index out of range");
    const void *Ea, *Eb, *Ha, *Hb;
    int64_t n1, n2;
    plane_fields(g, axis, &Ea, &Eb, &Ha, &Hb, &n1, &n2);
    double *part;
    int rc = partials(&part);
    if (rc) return rc;
    KernelTimer t("trace_plane");
    if (g->field_dtype == MP_F32) {
        trace_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)Ea, (const float *)Eb,
            axis, i_pre, i_post, mode_w, area, n1, n2, g->nx, g->ny, part);
    } else if (g->field_dtype == MP_F64) {
        trace_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)Ea, (const double *)Eb,
            axis, i_pre, i_post, mode_w, area, n1, n2, g->nx, g->ny, part);
    } else {
        return fail_arg("mp_probe_trace_plane: unsupported field dtype");
    }
    MP_LAUNCH_CHECK("trace_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, 4, 4, 4, 1.0, out);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

MP_API int mp_flux_plane(const mp_yee *g, int32_t axis, int64_t index, const double *area, double *out,
                         int64_t slot) {
    if (!g || !area || !out) return fail_arg("mp_flux_plane: null argument");
    if (axis < 0 || axis > 2) return fail_arg("3-D flux axis must be x, y or z");
    const int64_t nax = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    if (index < 0 || index >= nax || slot < 0) return fail_arg("mp_flux_plane: index out of range");
    const void *Ea, *Eb, *Ha, *Hb;
    int64_t n1, n2;
    plane_fields(g, axis, &Ea, &Eb, &Ha, &Hb, &n1, &n2);
    double *part;
    int rc = partials(&part);
    if (rc) return rc;
    KernelTimer t("flux_plane");
    if (g->field_dtype == MP_F32) {
        flux_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)Ea, (const float *)Eb,
            (const float *)Ha, (const float *)Hb, axis, index, area, n1, n2, g->nx, g->ny, part);
    } else if (g->field_dtype == MP_F64) {
        flux_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)Ea, (const double *)Eb,
            (const double *)Ha, (const double *)Hb, axis, index, area, n1, n2, g->nx, g->ny, part);
    } else {
        return fail_arg("mp_flux_plane: unsupported field dtype");
    }
    MP_LAUNCH_CHECK("flux_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, 1, 1, 1, 1.0, out + slot);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

namespace mp {

static double *g_shot_part = nullptr;  // SHOT_NV * SHOT_BLOCKS * (nb + 1) doubles
static int64_t g_shot_cap = 0;

static int shot_partials(int nb, double **out) {
    std::lock_guard<std::mutex> lk(g_red_mu);
    const int64_t need = (int64_t)SHOT_NV * SHOT_BLOCKS * (nb + 1);
    if (need > g_shot_cap) {
        if (g_shot_part) MP_HIP(hipFree(g_shot_part));
        g_shot_part = nullptr;
        g_shot_cap = 0;
        MP_HIP(hipMalloc((void **)&g_shot_part, sizeof(double) * need));
        g_shot_cap = need;
    }
    *out = g_shot_part;
    return MP_OK;
}

static int shot_plane(const mp_plane *p, ShotPlane &s) {
    if (!p->dft_ea || !p->dft_eb || !p->dft_ha || !p->dft_hb || !p->energy)
        return fail_arg("mp_probe_shot_reduce: null plane buffer");
    s.ea = (const double2 *)p->dft_ea;
    s.eb = (const double2 *)p->dft_eb;
    s.ha = (const double2 *)p->dft_ha;
    s.hb = (const double2 *)p->dft_hb;
    s.energy = p->energy;
    return MP_OK;
}

}  // namespace mp

namespace mp {

// Argument checks shared by the two shot-reduction entry points.
static int shot_args(const char *who, const mp_plane *pre, const mp_plane *post, int64_t n1, int64_t n2, int32_t nb,
                     int32_t center_bin, const double *mode_w, const double *area, const double *out, ShotPlane &a,
                     ShotPlane &b) {
    if (!pre || !post || !mode_w || !area || !out) return fail_arg("%s: null argument", who);
    if (n1 < 1 || n2 < 1) return fail_arg("%s: plane sizes must be positive", who);
    if (nb < 1 || nb > 65535) return fail_arg("%s: nb must be in [1, 65535]", who);
    if (center_bin < 0 || center_bin >= nb) return fail_arg("%s: center_bin out of range", who);
    if (int rc = shot_plane(pre, a)) return rc;
    return shot_plane(post, b);
}

// The two launches of a shot reduction; part holds SHOT_NV * SHOT_BLOCKS * (nb + 1)
// doubles, every one of which the first launch writes before the second reads it.
static int shot_launch(const ShotPlane &a, const ShotPlane &b, int64_t np, int32_t nb, int32_t center_bin,
                       const double *mode_w, const double *area, double *part, double *out) {
    KernelTimer t("probe_shot_reduce");
    shot_partial_kernel<<<dim3(SHOT_BLOCKS, nb + 1), RED_THREADS, 0, stream()>>>(a, b, mode_w, area, np, nb,
                                                                                   center_bin, part);
    MP_LAUNCH_CHECK("shot_partial_kernel");
    const int nout = 4 * nb + SHOT_NV;
    shot_finalize_kernel<<<(unsigned)cdiv(nout, 256), 256, 0, stream()>>>(part, nb, out);
    MP_LAUNCH_CHECK("shot_finalize_kernel");
    return MP_OK;
}

// Partials of mp_probe_shot_reduce_noalloc: a fixed device array, one copy per GPU, held
// until the process ends.  Host code never uses the name `scratch` as an address (in host
// code it is a stand-in, not device memory): it uses the address hipGetSymbolAddress gives
// for the current device, resolved once per device at mp_init or at the first call.
constexpr int SHOT_NOALLOC_DOUBLES = SHOT_NV * SHOT_BLOCKS * (MP_SHOT_NOALLOC_MAX_BINS + 1);
static __device__ __align__(256) double scratch[SHOT_NOALLOC_DOUBLES];

constexpr int MAX_DEVICES = 64;
static double *g_scratch_dev[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// The current device's copy of `scratch` (resolved with that device current).
int noalloc_scratch(double **out) {
    int dev = 0;
    MP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail_arg("no-allocation scratch: device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        MP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail_arg("no-allocation scratch: no device address on device %d", dev);
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return MP_OK;
}

}  // namespace mp

MP_API int mp_probe_shot_reduce(const mp_plane *pre, const mp_plane *post, int64_t n1, int64_t n2, int32_t nb,
                                int32_t center_bin, const double *mode_w, const double *area, double *out) {
    ShotPlane a, b;
    if (int rc = shot_args("mp_probe_shot_reduce", pre, post, n1, n2, nb, center_bin, mode_w, area, out, a, b))
        return rc;
    double *part;
    if (int rc = shot_partials(nb, &part)) return rc;
    return shot_launch(a, b, n1 * n2, nb, center_bin, mode_w, area, part, out);
}

MP_API int mp_probe_shot_reduce_noalloc(const mp_plane *pre, const mp_plane *post, int64_t n1, int64_t n2,
                                        int32_t nb, int32_t center_bin, const double *mode_w, const double *area,
                                        double *out) {
    ShotPlane a, b;
    if (int rc = shot_args("mp_probe_shot_reduce_noalloc", pre, post, n1, n2, nb, center_bin, mode_w, area, out, a,
                           b))
        return rc;
    if (nb > MP_SHOT_NOALLOC_MAX_BINS)
        return fail_arg("mp_probe_shot_reduce_noalloc: nb = %d exceeds the maximum of %d bins of the fixed scratch",
                        nb, MP_SHOT_NOALLOC_MAX_BINS);
    double *part;
    if (int rc = noalloc_scratch(&part)) return rc;
    return shot_launch(a, b, n1 * n2, nb, center_bin, mode_w, area, part, out);
}

MP_API int mp_plane_overlap(const mp_yee *g, int32_t component, int32_t axis, int64_t index, const double *weight,
                            double *out) {
    if (!g || !out) return fail_arg("mp_plane_overlap: null argument");
    if (component < 0 || component > 5) return fail_arg("mp_plane_overlap: component must be 0..5 (Ex..Hz)");
    if (axis < 0 || axis > 2) return fail_arg("mp_plane_overlap: bad axis");
    const int64_t nax = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    if (index < 0 || index >= nax) return fail_arg("mp_plane_overlap: plane index out of range");
    const void *comps[6] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz};
    const void *F = comps[component];
    if (!F) return fail_arg("mp_plane_overlap: null field");
    const int64_t n1 = axis == 0 ? g->ny : (axis == 1 ? g->nz : g->nx);
    const int64_t n2 = axis == 0 ? g->nz : (axis == 1 ? g->nx : g->ny);
    double *part;
    int rc = partials(&part);
    if (rc) return rc;
    KernelTimer t("plane_overlap");
    if (g->field_dtype == MP_F32) {
        overlap_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)F, axis, index, weight,
                                                                                n1, n2, g->nx, g->ny, part);
    } else if (g->field_dtype == MP_F64) {
        overlap_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)F, axis, index,
                                                                                 weight, n1, n2, g->nx, g->ny, part);
    } else {
        return fail_arg("mp_plane_overlap: unsupported field dtype");
    }
    MP_LAUNCH_CHECK("overlap_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, 1, 1, 1, 1.0, out);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

MP_API int mp_film_reduce(const mp_film *f, double *out) {
    if (!f || !out) return fail_arg("mp_film_reduce: null argument");
    if (f->n <= 0) return mp_fill(out, MP_F64, 10, 0.0);
    double *part;
    int rc = partials(&part);
    if (rc) return rc;
    KernelTimer t("film_reduce");
    if (f->dtype == MP_F32) {
        film_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)f->m_TM_x,
            (const float *)f->m_TM_y, (const float *)f->m_TM_z, (const float *)f->m_RE_x, (const float *)f->m_RE_y,
            (const float *)f->m_RE_z, (const float *)f->Te, (const float *)f->Tl, f->n, part);
    } else if (f->dtype == MP_F64) {
        film_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)f->m_TM_x,
            (const double *)f->m_TM_y, (const double *)f->m_TM_z, (const double *)f->m_RE_x, (const double *)f->m_RE_y,
            (const double *)f->m_RE_z, (const double *)f->Te, (const double *)f->Tl, f->n, part);
    } else {
        return fail_arg("mp_film_reduce: unsupported dtype");
    }
    MP_LAUNCH_CHECK("film_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, RED_MAXV, 9, 10, 1.0 / (double)f->n, out);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

MP_API int mp_film_average(const mp_film *f, double *out) {
    if (!f || !out) return fail_arg("mp_film_average: null argument");
    if (f->n <= 0) return mp_fill(out, MP_F64, 13, 0.0);
    double *part;
    int rc = partials(&part);
    if (rc) return rc;
    KernelTimer t("film_average");
    if (f->dtype == MP_F32) {
        film_average_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)f->m_TM_x,
            (const float *)f->m_TM_y, (const float *)f->m_TM_z, (const float *)f->m_RE_x, (const float *)f->m_RE_y,
            (const float *)f->m_RE_z, (const float *)f->Te, (const float *)f->Tl, (const float *)f->Ts_TM,
            (const float *)f->Ts_RE, f->n, part);
    } else if (f->dtype == MP_F64) {
        film_average_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)f->m_TM_x,
            (const double *)f->m_TM_y, (const double *)f->m_TM_z, (const double *)f->m_RE_x, (const double *)f->m_RE_y,
            (const double *)f->m_RE_z, (const double *)f->Te, (const double *)f->Tl, (const double *)f->Ts_TM,
            (const double *)f->Ts_RE, f->n, part);
    } else {
        return fail_arg("mp_film_average: unsupported film dtype");
    }
    MP_LAUNCH_CHECK("film_average_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, RED_MAXV, 13, 13, 1.0, out);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

MP_API int mp_switching_metrics(const mp_film *f, double core_fraction, double *out) {
    if (!f || !out) return fail_arg("mp_switching_metrics: null argument");
    if (f->n <= 0) return fail_arg("mp_switching_metrics: empty film");
    if (!f->U_abs) return fail_arg("mp_switching_metrics: null U_abs");
    double *part;
    int rc = partials(&part);
    if (rc) return rc;
    KernelTimer t("switching_metrics");
    if (f->dtype == MP_F32) {
        switch_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)f->m_TM_x,
            (const float *)f->m_RE_x, f->U_abs, f->n, part);
    } else if (f->dtype == MP_F64) {
        switch_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)f->m_TM_x,
            (const double *)f->m_RE_x, f->U_abs, f->n, part);
    } else {
        return fail_arg("mp_switching_metrics: unsupported film dtype");
    }
    MP_LAUNCH_CHECK("switch_partial_kernel");
    switch_finalize_kernel<<<1, 64, 0, stream()>>>(part, f->n, out);
    MP_LAUNCH_CHECK("switch_finalize_kernel");
    if (f->dtype == MP_F32) {
        core_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)f->m_TM_x,
            (const float *)f->m_RE_x, f->U_abs, f->n, out, core_fraction, part);
    } else {
        core_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)f->m_TM_x,
            (const double *)f->m_RE_x, f->U_abs, f->n, out, core_fraction, part);
    }
    MP_LAUNCH_CHECK("core_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, RED_MAXV, 4, 4, 1.0, out + 9);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

MP_API int mp_any_nonfinite(const void *x, int32_t dtype, int64_t n, int32_t *flag) {
    if (!flag) return fail_arg("mp_any_nonfinite: null flag");
    if (n <= 0) return MP_OK;
    if (!x) return fail_arg("mp_any_nonfinite: null array");
    const unsigned blocks = (unsigned)std::min<int64_t>(cdiv(n, 256), 4096);
    KernelTimer t("nonfinite");
    if (dtype == MP_F32) nonfinite_kernel<float><<<blocks, 256, 0, stream()>>>((const float *)x, n, flag);
    else if (dtype == MP_F64) nonfinite_kernel<double><<<blocks, 256, 0, stream()>>>((const double *)x, n, flag);
    else return fail_arg("mp_any_nonfinite: unsupported dtype");
    MP_LAUNCH_CHECK("nonfinite_kernel");
    return MP_OK;
}

// ---- generic device reductions and plane copies behind the backend hooks ---------------------
// reduce_sum / reduce_mean / reduce_count_negative / reduce_norm_mean (src/core/Backend.jl:
// 42-55), sum(x; dims) over leading dimensions (src/sim/ProbeReadout.jl:291) and
// plane_to_host (Backend.jl:61-83) on device arrays, which the reference computes with
// GPUArrays mapreduce and views.  Fixed partitions and ordered final sums: a call gives
// the same result every time.
namespace mp {

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void reduce_partial_kernel(const T *x, const T *y, const T *z, int64_t n,
                                                                     int op, double *part) {
    double v[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS) {
        const double a = (double)x[i];
        if (op == MP_RED_SUM) {
            v[0] += a;
        } else if (op == MP_RED_COUNT_NEGATIVE) {
            v[0] += x[i] < (T)0 ? 1.0 : 0.0;
        } else {
            const double b = (double)y[i], c = (double)z[i];
            v[0] += sqrt(a * a + b * b + c * c);
        }
    }
    block_sum_store<1>(v, part + blockIdx.x);
}

// out[b] = sum of block b (block elements, consecutive), one workgroup per block.
__global__ __launch_bounds__(RED_THREADS) void sum_blocks_kernel(const double *x, int64_t block, double *out) {
    double v[1] = {0.0};
    const double *xb = x + (int64_t)blockIdx.x * block;
    for (int64_t i = threadIdx.x; i < block; i += RED_THREADS) v[0] += xb[i];
    block_sum_store<1>(v, out + blockIdx.x);
}

// The plane `index` normal to `axis` in the order of the reference's views: x -> field[i,:,:]
// (ny, nz), y -> field[:,i,:] (nx, nz), z -> field[:,:,i] (nx, ny), first index fastest.
template <typename T>
__global__ __launch_bounds__(256) void plane_gather_kernel(const T *f, int axis, int64_t idx, int64_t nx, int64_t ny,
                                                           int64_t n1, int64_t n2, T *out) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n1 * n2) return;
    const int64_t a = c % n1, b = c / n1;
    int64_t o;
    if (axis == 0) o = idx + nx * (a + ny * b);
    else if (axis == 1) o = a + nx * (idx + ny * b);
    else o = a + nx * (b + ny * idx);
    out[c] = f[o];
}

// probe_trace_plane_kernel! (Kernels.jl:467-478) on x-normal planes i_pre, i_post:
// work (ny, nz, 4) = Ey_pre*w, Ez_pre*w, Ey_post*w, Ez_post*w with w = mode_w*area.
template <typename S>
__global__ __launch_bounds__(256) void trace_samples_kernel(const S *Ey, const S *Ez, int64_t ipre, int64_t ipost,
                                                            const double *mode_w, const double *area, int64_t nx,
                                                            int64_t ny, int64_t nz, double *work) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t np = ny * nz;
    if (c >= np) return;
    const double w = mode_w[c] * area[c];
    const int64_t o1 = ipre + nx * c, o2 = ipost + nx * c;
    work[c] = (double)Ey[o1] * w;
    work[c + np] = (double)Ez[o1] * w;
    work[c + 2 * np] = (double)Ey[o2] * w;
    work[c + 3 * np] = (double)Ez[o2] * w;
}

// One probe movie frame (_record_probe_frame!, src/sim/ProbeReadout.jl:193-215): |Ey| and
// |Ez| rounded to float on the z-plane kc (nx, ny) and on the x-planes ipre, ipost (ny, nz),
// first index fastest.  blockIdx.y picks the plane (0: z, 1: pre, 2: post).
template <typename S>
__global__ __launch_bounds__(256) void probe_frames_kernel(const S *Ey, const S *Ez, int64_t nx, int64_t ny,
                                                           int64_t nz, int64_t kc, int64_t ipre, int64_t ipost,
                                                           float *xy_ey, float *xy_ez, float *pre_ey, float *pre_ez,
                                                           float *post_ey, float *post_ez) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.y == 0) {
        if (c >= nx * ny) return;
        const int64_t o = c + nx * ny * kc;
        xy_ey[c] = (float)__builtin_fabs(Ey[o]);
        xy_ez[c] = (float)__builtin_fabs(Ez[o]);
    } else {
        if (c >= ny * nz) return;
        const int64_t o = (blockIdx.y == 1 ? ipre : ipost) + nx * c;
        float *dy = blockIdx.y == 1 ? pre_ey : post_ey;
        float *dz = blockIdx.y == 1 ? pre_ez : post_ez;
        dy[c] = (float)__builtin_fabs(Ey[o]);
        dz[c] = (float)__builtin_fabs(Ez[o]);
    }
}

}  // namespace mp

MP_API int mp_reduce(const void *x, const void *y, const void *z, int32_t dtype, int64_t n, int32_t op, double *out) {
    if (!out) return fail_arg("mp_reduce: null output");
    if (op != MP_RED_SUM && op != MP_RED_COUNT_NEGATIVE && op != MP_RED_NORM_SUM)
        return fail_arg("mp_reduce: unknown op %d", op);
    if (n <= 0) return mp_fill(out, MP_F64, 1, 0.0);
    if (!x || (op == MP_RED_NORM_SUM && (!y || !z))) return fail_arg("mp_reduce: null input");
    double *part;
    if (int rc = partials(&part)) return rc;
    KernelTimer t("reduce");
    if (dtype == MP_F32)
        reduce_partial_kernel<float><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const float *)x, (const float *)y,
                                                                              (const float *)z, n, op, part);
    else if (dtype == MP_F64)
        reduce_partial_kernel<double><<<RED_BLOCKS, RED_THREADS, 0, stream()>>>((const double *)x, (const double *)y,
                                                                               (const double *)z, n, op, part);
    else
        return fail_arg("mp_reduce: unsupported dtype %d", dtype);
    MP_LAUNCH_CHECK("reduce_partial_kernel");
    finalize_kernel<<<1, 64, 0, stream()>>>(part, 1, 1, 1, 1.0, out);
    MP_LAUNCH_CHECK("finalize_kernel");
    return MP_OK;
}

MP_API int mp_sum_blocks(const double *x, int64_t block, int64_t nblocks, double *out) {
    if (!x || !out) return fail_arg("mp_sum_blocks: null argument");
    if (block < 0 || nblocks < 0 || nblocks > 65535) return fail_arg("mp_sum_blocks: bad sizes");
    if (nblocks == 0) return MP_OK;
    if (block == 0) return mp_fill(out, MP_F64, nblocks, 0.0);
    KernelTimer t("sum_blocks");
    sum_blocks_kernel<<<(unsigned)nblocks, RED_THREADS, 0, stream()>>>(x, block, out);
    MP_LAUNCH_CHECK("sum_blocks_kernel");
    return MP_OK;
}

MP_API int mp_plane_gather(const void *field, int32_t dtype, int64_t nx, int64_t ny, int64_t nz, int32_t axis,
                           int64_t index, void *out) {
    if (!field || !out) return fail_arg("mp_plane_gather: null argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail_arg("mp_plane_gather: bad dims");
    if (axis < 0 || axis > 2) return fail_arg("3-D field plane axis must be :x, :y, or :z");
    const int64_t nax = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (index < 0 || index >= nax) return fail_arg("mp_plane_gather: plane index out of range");
    const int64_t n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
    KernelTimer t("plane_gather");
    if (dtype == MP_F32)
        plane_gather_kernel<float><<<cdiv(n1 * n2, 256), 256, 0, stream()>>>((const float *)field, axis, index, nx, ny,
                                                                             n1, n2, (float *)out);
    else if (dtype == MP_F64)
        plane_gather_kernel<double><<<cdiv(n1 * n2, 256), 256, 0, stream()>>>((const double *)field, axis, index, nx,
                                                                              ny, n1, n2, (double *)out);
    else
        return fail_arg("mp_plane_gather: unsupported dtype %d", dtype);
    MP_LAUNCH_CHECK("plane_gather_kernel");
    return MP_OK;
}

MP_API int mp_probe_trace_samples(const mp_yee *g, int64_t i_pre, int64_t i_post, const double *mode_w,
                                  const double *area, double *work) {
    if (!g || !mode_w || !area || !work) return fail_arg("mp_probe_trace_samples: null argument");
    if (i_pre < 0 || i_pre >= g->nx || i_post < 0 || i_post >= g->nx)
        return fail_arg("mp_probe_trace_samples: plane index out of range");
    const int64_t np = g->ny * g->nz;
    if (np == 0) return MP_OK;
    KernelTimer t("trace_samples");
    if (g->field_dtype == MP_F32)
        trace_samples_kernel<float><<<cdiv(np, 256), 256, 0, stream()>>>(
            (const float *)g->ey, (const float *)g->ez, i_pre, i_post, mode_w, area, g->nx, g->ny, g->nz, work);
    else if (g->field_dtype == MP_F64)
        trace_samples_kernel<double><<<cdiv(np, 256), 256, 0, stream()>>>(
            (const double *)g->ey, (const double *)g->ez, i_pre, i_post, mode_w, area, g->nx, g->ny, g->nz, work);
    else
        return fail_arg("mp_probe_trace_samples: unsupported field dtype");
    MP_LAUNCH_CHECK("trace_samples_kernel");
    return MP_OK;
}

MP_API int mp_probe_frames(const mp_yee *g, int64_t kc, int64_t i_pre, int64_t i_post, float *xy_ey, float *xy_ez,
                           float *yz_ey_pre, float *yz_ez_pre, float *yz_ey_post, float *yz_ez_post) {
    if (!g || !xy_ey || !xy_ez || !yz_ey_pre || !yz_ez_pre || !yz_ey_post || !yz_ez_post)
        return fail_arg("mp_probe_frames: null argument");
    if (kc < 0 || kc >= g->nz) return fail_arg("mp_probe_frames: z slice index out of range");
    if (i_pre < 0 || i_pre >= g->nx || i_post < 0 || i_post >= g->nx)
        return fail_arg("mp_probe_frames: plane index out of range");
    const int64_t n = std::max(g->nx * g->ny, g->ny * g->nz);
    if (n == 0) return MP_OK;
    const dim3 grid(cdiv(n, 256), 3);
    KernelTimer t("probe_frames");
    if (g->field_dtype == MP_F32)
        probe_frames_kernel<float><<<grid, 256, 0, stream()>>>((const float *)g->ey, (const float *)g->ez, g->nx, g->ny,
                                                               g->nz, kc, i_pre, i_post, xy_ey, xy_ez, yz_ey_pre,
                                                               yz_ez_pre, yz_ey_post, yz_ez_post);
    else if (g->field_dtype == MP_F64)
        probe_frames_kernel<double><<<grid, 256, 0, stream()>>>((const double *)g->ey, (const double *)g->ez, g->nx,
                                                                g->ny, g->nz, kc, i_pre, i_post, xy_ey, xy_ez,
                                                                yz_ey_pre, yz_ez_pre, yz_ey_post, yz_ez_post);
    else
        return fail_arg("mp_probe_frames: unsupported field dtype");
    MP_LAUNCH_CHECK("probe_frames_kernel");
    return MP_OK;
}
