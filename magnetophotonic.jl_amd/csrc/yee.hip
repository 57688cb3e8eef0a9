// Yee H / E leapfrog with CFS-CPML convolution memory, FP32 (or FP64) on gfx950.
//
// Reference: src/fdtd/Kernels.jl:9-152 (yee_H_3d_kernel!, yee_E_3d_kernel!,
// _cpml_term_{x,y,z}!, _pml_slab) and the launchers :522-565.
//
// Layout: volumes column-major, x fastest.  Three kernel families:
//  - FP32 with rows of a multiple of 4 cells (the vector form): each lane owns a float4
//    group of 4 x-consecutive cells.  The interior box (no CPML slab, no boundary cell)
//    is updated by lean kernels without psi terms; the CPML shell around it by the
//    general vector kernels, whose lanes decode their group from the shell layout and
//    load every psi value before the cell loop.  Where the box is empty the general
//    kernels cover the whole volume, a z-plane per workgroup range.
//  - Other shapes and FP64: the scalar kernels.  A workgroup is a 64 x 4 tile of (i, j)
//    columns (one wave64 per x-row, so every global access of a wave is 64 consecutive
//    elements); each thread marches a chunk of z planes and keeps the previous plane's
//    values in registers, so the z-neighbour never comes from memory twice.
// CPML psi slabs are touched only by lanes inside a slab.
//
// Arithmetic order is the reference's: derivative = (E[n] - E[c]) * inv_d, psi update
// = fma(b, psi, a*deriv) (Julia muladd), curl terms and the D/H updates exactly as
// written there, compiled with -ffp-contract=off.
#include <algorithm>
#include <cstdlib>

#include "common.h"
#include "yee.h"

namespace mp {

constexpr int TX = 64;
constexpr int TY = 4;

template <typename C, typename P> struct PsiIO {
    static __device__ __forceinline__ C load(const P *p, int64_t o, float) { return (C)p[o]; }
    static __device__ __forceinline__ void store(P *p, int64_t o, C v, float) { p[o] = (P)v; }
};
template <> struct PsiIO<float, bf16> {
    static __device__ __forceinline__ float load(const bf16 *p, int64_t o, float inv_sc) { return ld<float>(p[o]) * inv_sc; }
    static __device__ __forceinline__ void store(bf16 *p, int64_t o, float v, float sc) { p[o] = st<bf16>(v * sc); }
};
template <> struct PsiIO<float, f16s> {
    static __device__ __forceinline__ float load(const f16s *p, int64_t o, float inv_sc) { return ld<float>(p[o]) * inv_sc; }
    static __device__ __forceinline__ void store(f16s *p, int64_t o, float v, float sc) { p[o] = st<f16s>(v * sc); }
};

// _pml_slab (Kernels.jl:9-14), 0-based: region 0 none, 1 lo, 2 hi; *js slab coordinate.
__device__ __forceinline__ int pml_slab(int64_t idx, int64_t N, int64_t w, int64_t *js) {
    if (w <= 0) { *js = 0; return 0; }
    if (idx < w) { *js = idx; return 1; }
    if (idx >= N - w) { *js = idx - (N - w); return 2; }
    *js = 0;
    return 0;
}

// _cpml_term_*! (Kernels.jl:16-56)
template <typename C, typename P>
__device__ __forceinline__ C cpml_term(P *lo, P *hi, int region, int64_t off, C b, C a, C deriv,
                                       float sc, float isc) {
    if (region == 1) {
        C p = fmaT<C>(b, PsiIO<C, P>::load(lo, off, isc), a * deriv);
        PsiIO<C, P>::store(lo, off, p, sc);
        return p;
    } else if (region == 2) {
        C p = fmaT<C>(b, PsiIO<C, P>::load(hi, off, isc), a * deriv);
        PsiIO<C, P>::store(hi, off, p, sc);
        return p;
    }
    return (C)0;
}

struct CpmlArgs {
    const void *ik[3], *a[3], *b[3];
    void *psi[12];
    int64_t w[3];
    float sc, isc;
};

template <typename S, typename C, typename P>
__global__ __launch_bounds__(TX *TY) void yee_H_kernel(
    const S *__restrict__ Ex, const S *__restrict__ Ey, const S *__restrict__ Ez,
    S *__restrict__ Hx, S *__restrict__ Hy, S *__restrict__ Hz,
    const C *__restrict__ idcx, const C *__restrict__ idcy, const C *__restrict__ idcz,
    CpmlArgs cp, int64_t nx, int64_t ny, int64_t nz, int64_t kchunk, C dt_mu0) {
    const int64_t i = (int64_t)blockIdx.x * TX + threadIdx.x;
    const int64_t j = (int64_t)blockIdx.y * TY + threadIdx.y;
    if (i >= nx - 1 || j >= ny - 1) return;
    const int64_t k0 = (int64_t)blockIdx.z * kchunk;
    const int64_t k1 = min(k0 + kchunk, nz - 1);
    if (k0 >= k1) return;
    const int64_t sy = nx, sz = nx * ny;
    const C *xk = (const C *)cp.ik[0], *xa = (const C *)cp.a[0], *xb = (const C *)cp.b[0];
    const C *yk = (const C *)cp.ik[1], *ya = (const C *)cp.a[1], *yb = (const C *)cp.b[1];
    const C *zk = (const C *)cp.ik[2], *za = (const C *)cp.a[2], *zb = (const C *)cp.b[2];
    P **ps = (P **)cp.psi;
    const int64_t wx = cp.w[0], wy = cp.w[1], wz = cp.w[2];

    int64_t jsx, jsy;
    const int rx = pml_slab(i, nx, wx, &jsx);
    const int ry = pml_slab(j, ny, wy, &jsy);
    const C idx_ = idcx[i], idy_ = idcy[j];
    const C xk_ = xk[i], yk_ = yk[j];
    const C xa_ = xa[i], xb_ = xb[i], ya_ = ya[j], yb_ = yb[j];

    int64_t o = i + sy * j + sz * k0;
    C ex_c = (C)Ex[o], ey_c = (C)Ey[o];
    for (int64_t k = k0; k < k1; ++k, o += sz) {
        const C ex_n = (C)Ex[o + sz], ey_n = (C)Ey[o + sz];
        const C ez_c = (C)Ez[o];
        const C idz_ = idcz[k];
        const C dEz_dy = ((C)Ez[o + sy] - ez_c) * idy_;
        const C dEy_dz = (ey_n - ey_c) * idz_;
        const C dEx_dz = (ex_n - ex_c) * idz_;
        const C dEz_dx = ((C)Ez[o + 1] - ez_c) * idx_;
        const C dEy_dx = ((C)Ey[o + 1] - ey_c) * idx_;
        const C dEx_dy = ((C)Ex[o + sy] - ex_c) * idy_;

        int64_t jsz;
        const int rz = pml_slab(k, nz, wz, &jsz);
        const int64_t ox = jsx + wx * (j + ny * k);
        const int64_t oy = i + nx * (jsy + wy * k);
        const int64_t oz = i + nx * (j + ny * jsz);
        const C zk_ = zk[k];
        const C za_ = rz ? za[k] : (C)0, zb_ = rz ? zb[k] : (C)0;

        const C pHxy = cpml_term<C, P>(ps[0], ps[1], ry, oy, yb_, ya_, dEz_dy, cp.sc, cp.isc);
        const C pHxz = cpml_term<C, P>(ps[2], ps[3], rz, oz, zb_, za_, dEy_dz, cp.sc, cp.isc);
        Hx[o] = (S)((C)Hx[o] - dt_mu0 * ((dEz_dy * yk_ + pHxy) - (dEy_dz * zk_ + pHxz)));

        const C pHyz = cpml_term<C, P>(ps[4], ps[5], rz, oz, zb_, za_, dEx_dz, cp.sc, cp.isc);
        const C pHyx = cpml_term<C, P>(ps[6], ps[7], rx, ox, xb_, xa_, dEz_dx, cp.sc, cp.isc);
        Hy[o] = (S)((C)Hy[o] - dt_mu0 * ((dEx_dz * zk_ + pHyz) - (dEz_dx * xk_ + pHyx)));

        const C pHzx = cpml_term<C, P>(ps[8], ps[9], rx, ox, xb_, xa_, dEy_dx, cp.sc, cp.isc);
        const C pHzy = cpml_term<C, P>(ps[10], ps[11], ry, oy, yb_, ya_, dEx_dy, cp.sc, cp.isc);
        Hz[o] = (S)((C)Hz[o] - dt_mu0 * ((dEy_dx * xk_ + pHzx) - (dEx_dy * yk_ + pHzy)));

        ex_c = ex_n;
        ey_c = ey_n;
    }
}

template <typename S, typename C, typename P>
__global__ __launch_bounds__(TX *TY) void yee_E_kernel(
    S *__restrict__ Ex, S *__restrict__ Ey, S *__restrict__ Ez,
    const S *__restrict__ Hx, const S *__restrict__ Hy, const S *__restrict__ Hz,
    S *__restrict__ Dx, S *__restrict__ Dy, S *__restrict__ Dz,
    const S *__restrict__ iex, const S *__restrict__ iey, const S *__restrict__ iez,
    const C *__restrict__ iddx, const C *__restrict__ iddy, const C *__restrict__ iddz,
    CpmlArgs cp, int64_t nx, int64_t ny, int64_t nz, int64_t kchunk, C dt) {
    const int64_t i = (int64_t)blockIdx.x * TX + threadIdx.x;
    const int64_t j = (int64_t)blockIdx.y * TY + threadIdx.y;
    if (i < 1 || j < 1 || i >= nx || j >= ny) return;
    const int64_t k0 = max((int64_t)blockIdx.z * kchunk, (int64_t)1);
    const int64_t k1 = min((int64_t)blockIdx.z * kchunk + kchunk, nz);
    if (k0 >= k1) return;
    const int64_t sy = nx, sz = nx * ny;
    const C *xk = (const C *)cp.ik[0], *xa = (const C *)cp.a[0], *xb = (const C *)cp.b[0];
    const C *yk = (const C *)cp.ik[1], *ya = (const C *)cp.a[1], *yb = (const C *)cp.b[1];
    const C *zk = (const C *)cp.ik[2], *za = (const C *)cp.a[2], *zb = (const C *)cp.b[2];
    P **ps = (P **)cp.psi;
    const int64_t wx = cp.w[0], wy = cp.w[1], wz = cp.w[2];

    int64_t jsx, jsy;
    const int rx = pml_slab(i, nx, wx, &jsx);
    const int ry = pml_slab(j, ny, wy, &jsy);
    const C idx_ = iddx[i], idy_ = iddy[j];
    const C xk_ = xk[i], yk_ = yk[j];
    const C xa_ = xa[i], xb_ = xb[i], ya_ = ya[j], yb_ = yb[j];

    int64_t o = i + sy * j + sz * k0;
    C hx_p = (C)Hx[o - sz], hy_p = (C)Hy[o - sz];
    for (int64_t k = k0; k < k1; ++k, o += sz) {
        const C hx_c = (C)Hx[o], hy_c = (C)Hy[o], hz_c = (C)Hz[o];
        const C idz_ = iddz[k];
        const C dHz_dy = (hz_c - (C)Hz[o - sy]) * idy_;
        const C dHy_dz = (hy_c - hy_p) * idz_;
        const C dHx_dz = (hx_c - hx_p) * idz_;
        const C dHz_dx = (hz_c - (C)Hz[o - 1]) * idx_;
        const C dHy_dx = (hy_c - (C)Hy[o - 1]) * idx_;
        const C dHx_dy = (hx_c - (C)Hx[o - sy]) * idy_;

        int64_t jsz;
        const int rz = pml_slab(k, nz, wz, &jsz);
        const int64_t ox = jsx + wx * (j + ny * k);
        const int64_t oy = i + nx * (jsy + wy * k);
        const int64_t oz = i + nx * (j + ny * jsz);
        const C zk_ = zk[k];
        const C za_ = rz ? za[k] : (C)0, zb_ = rz ? zb[k] : (C)0;

        const C pDxy = cpml_term<C, P>(ps[0], ps[1], ry, oy, yb_, ya_, dHz_dy, cp.sc, cp.isc);
        const C pDxz = cpml_term<C, P>(ps[2], ps[3], rz, oz, zb_, za_, dHy_dz, cp.sc, cp.isc);
        const C cx_term = (dHz_dy * yk_ + pDxy) - (dHy_dz * zk_ + pDxz);

        const C pDyz = cpml_term<C, P>(ps[4], ps[5], rz, oz, zb_, za_, dHx_dz, cp.sc, cp.isc);
        const C pDyx = cpml_term<C, P>(ps[6], ps[7], rx, ox, xb_, xa_, dHz_dx, cp.sc, cp.isc);
        const C cy_term = (dHx_dz * zk_ + pDyz) - (dHz_dx * xk_ + pDyx);

        const C pDzx = cpml_term<C, P>(ps[8], ps[9], rx, ox, xb_, xa_, dHy_dx, cp.sc, cp.isc);
        const C pDzy = cpml_term<C, P>(ps[10], ps[11], ry, oy, yb_, ya_, dHx_dy, cp.sc, cp.isc);
        const C cz_term = (dHy_dx * xk_ + pDzx) - (dHx_dy * yk_ + pDzy);

        const C dx_new = (C)Dx[o] + dt * cx_term;
        const C dy_new = (C)Dy[o] + dt * cy_term;
        const C dz_new = (C)Dz[o] + dt * cz_term;
        Dx[o] = (S)dx_new;
        Dy[o] = (S)dy_new;
        Dz[o] = (S)dz_new;
        Ex[o] = (S)(dx_new * (C)iex[o]);
        Ey[o] = (S)(dy_new * (C)iey[o]);
        Ez[o] = (S)(dz_new * (C)iez[o]);

        hx_p = hx_c;
        hy_p = hy_c;
    }
}

// ---- FP32 vector form: 4 consecutive x cells per lane ---------------------------------
// Each lane moves 16 B per access (dwordx4), which is what lets the stencil approach the
// HBM copy rate; a wave covers 256 cells of one x row.  The x neighbour one past the
// lane's group comes from the next lane by a cross-lane shuffle (lane 63 reads it from
// memory).  Cells the reference does not update (H at i = nx-1, E at i = 0) are written
// back unchanged.  Per-cell arithmetic is identical to the scalar kernels above.
constexpr int VX = 4;

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }
// Streamed-once arrays bypass cache retention so the L2 keeps the stencil's halo rows.
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4nt(const float *p) {
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4nt(float *p, float4 v) {
    v4f w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<v4f *>(p));
}

// Flattened plane mapping of the vector kernels: consecutive rows of a z-plane are
// adjacent in memory, so a plane is one run of nx*ny/4 float4 groups, row after row.
// Lanes take consecutive groups across row ends (no idle lanes when a row is not a
// multiple of 256 cells) and every load of a wave stays contiguous.
struct PlaneGroup { int64_t i0, j, bz; bool in_plane; };
__device__ __forceinline__ PlaneGroup plane_group(int64_t nx, int64_t ny, int64_t bpp, int tyv) {
    const int64_t t = blockIdx.x;
    const int64_t G = nx / VX;
    const int64_t g = (t % bpp) * (64 * tyv) + threadIdx.y * 64 + threadIdx.x;
    PlaneGroup r;
    r.bz = t / bpp;
    r.j = g / G;
    r.i0 = (g % G) * VX;
    r.in_plane = g < G * ny;
    return r;
}
__device__ __forceinline__ float &el(float4 &v, int e) { return (&v.x)[e]; }
__device__ __forceinline__ float elc(const float4 &v, int e) { return (&v.x)[e]; }

// psi of the four cells of a group in a y or z slab.  The slab coordinate of those slabs
// depends on j or k only, so the four values sit side by side (at i0 + nx * (...), a
// multiple of four) and move as one vector access instead of four strided ones.  A cell
// that the update skips keeps its stored bits: they are written back unchanged.
template <typename P> struct alignas(4 * sizeof(P)) Psi4 { P v[4]; };
template <typename P> __device__ __forceinline__ Psi4<P> psi4_load(const P *p, int64_t o) {
    return *reinterpret_cast<const Psi4<P> *>(p + o);
}
template <typename P> __device__ __forceinline__ void psi4_store(P *p, int64_t o, const Psi4<P> &q) {
    *reinterpret_cast<Psi4<P> *>(p + o) = q;
}
__device__ __forceinline__ float psi_val(float x, float) { return x; }
__device__ __forceinline__ float psi_val(bf16 x, float isc) { return ld<float>(x) * isc; }
__device__ __forceinline__ float psi_val(f16s x, float isc) { return ld<float>(x) * isc; }
__device__ __forceinline__ void psi_put(float &x, float v, float) { x = v; }
__device__ __forceinline__ void psi_put(bf16 &x, float v, float sc) { x = st<bf16>(v * sc); }
__device__ __forceinline__ void psi_put(f16s &x, float v, float sc) { x = st<f16s>(v * sc); }
// _cpml_term_*! (Kernels.jl:16-56) on element e of a loaded quad: psi = fma(b, psi, a*deriv)
template <typename P>
__device__ __forceinline__ float psi4_step(Psi4<P> &q, int e, float b, float a, float deriv, float sc, float isc) {
    const float p = fmaT<float>(b, psi_val(q.v[e], isc), a * deriv);
    psi_put(q.v[e], p, sc);
    return p;
}

// x-slab psi of the group's cells (arrays 6/7: the y component's x term, 8/9: the z
// component's), loaded before the update so that no load waits inside a per-cell branch.  A
// cell outside the slab reads the row's first slab entry (a valid address) and stores
// nothing.
template <typename P> struct PsiX {
    int r[VX];
    int64_t off[VX];
    P q6[VX], q8[VX];
    float a[VX], b[VX];
};
template <typename P>
__device__ __forceinline__ void psix_load(PsiX<P> &x, P *const *ps, bool load, bool xslab, int64_t ic, int64_t nx,
                                          int64_t wx, int64_t row, const float *xa, const float *xb) {
#pragma unroll
    for (int e = 0; e < VX; ++e) {
        int64_t jsx = 0;
        x.r[e] = xslab ? pml_slab(ic + e, nx, wx, &jsx) : 0;
        x.off[e] = jsx + wx * row;
        x.q6[e] = P{};
        x.q8[e] = P{};
        x.a[e] = 0.f;
        x.b[e] = 0.f;
    }
    if (load && xslab) {
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            x.q6[e] = ps[x.r[e] == 2 ? 7 : 6][x.off[e]];
            x.q8[e] = ps[x.r[e] == 2 ? 9 : 8][x.off[e]];
            x.a[e] = xa[ic + e];
            x.b[e] = xb[ic + e];
        }
    }
}
// _cpml_term_x! (Kernels.jl:16-56) on cell e: psi = fma(b, psi, a*deriv), kept in registers
template <typename P>
__device__ __forceinline__ float psix_step(P &q, float b, float a, float deriv, float sc, float isc) {
    const float p = fmaT<float>(b, psi_val(q, isc), a * deriv);
    psi_put(q, p, sc);
    return p;
}
// Stores of the x-slab psi after the cell loop: a store inside the loop would make every
// later wait for a load wait for the store as well (loads and stores share one counter).
// A skipped cell writes back the bits it loaded.
template <typename P> __device__ __forceinline__ void psix_store(const PsiX<P> &x, P *const *ps, bool store) {
    if (!store) return;
#pragma unroll
    for (int e = 0; e < VX; ++e) {
        if (x.r[e]) {
            ps[x.r[e] == 2 ? 7 : 6][x.off[e]] = x.q6[e];
            ps[x.r[e] == 2 ? 9 : 8][x.off[e]] = x.q8[e];
        }
    }
}

// Layout of the CPML shell around the interior box, in memory order: the mz low z planes
// whole; then in each of the nk box planes the my low rows whole, the x strips of the nj box
// rows (groups [0, gi0) and [gi0+ngi, G)) and the my high rows whole; then the mz high z
// planes whole.  Shell entry t is decoded arithmetically from these counts, so a lane's
// first loads depend on no load of its own.  Divisions by the uniform counts use the
// rounded-up reciprocal: q = umulhi(n, ceil(2^32/d)) is n/d or one more for n < 2^32.
struct ShellGeom {
    uint32_t n;                          // shell groups (0: plane mapping instead)
    uint32_t G, ny, nz, gi0, ngi, my, nj, mz;
    uint32_t pfull, pint, gx, nlo, nmid;   // groups per whole plane / per box plane, strip groups per row
    uint32_t r_pfull, r_pint, r_g, r_gx;   // ceil(2^32 / d) of those divisors
    uint32_t sgx;                        // strip groups per box row in this layout: gx, or 0 when the
                                         // x strips have a launch of their own
    uint32_t nstrip, r_nj;               // groups of all x strips; ceil(2^32 / nj)
    // the rest rows (every row outside the box's y or z range) split by x: the groups in the
    // box's x range (no x-slab cell) and the x edges; nnox = 0 when that split is not made
    uint32_t nrlo, nrmid, nnox, nedge;   // rows of the low z planes / of the box planes; groups
    uint32_t r_ny, r_2my, r_ngi;         // ceil(2^32 / d) of ny, 2 my, ngi
};
__device__ __forceinline__ uint32_t udiv(uint32_t n, uint32_t d, uint32_t r) {
    uint32_t q = __umulhi(n, r);
    if ((uint64_t)q * d > n) --q;
    return q;
}
__device__ __forceinline__ void shell_pos(const ShellGeom &sg, uint32_t t, uint32_t &g, uint32_t &j, uint32_t &k) {
    if (t < sg.nlo || t >= sg.nlo + sg.nmid) {   // a whole z plane
        const uint32_t u = t < sg.nlo ? t : t - sg.nlo - sg.nmid;
        const uint32_t kk = udiv(u, sg.pfull, sg.r_pfull), r = u - kk * sg.pfull;
        k = t < sg.nlo ? kk : sg.nz - sg.mz + kk;
        j = udiv(r, sg.G, sg.r_g);
        g = r - j * sg.G;
        return;
    }
    const uint32_t u = t - sg.nlo;
    const uint32_t kk = udiv(u, sg.pint, sg.r_pint);
    uint32_t r = u - kk * sg.pint;
    k = sg.mz + kk;
    const uint32_t lo = sg.my * sg.G, strips = sg.nj * sg.sgx;
    if (r < lo || r >= lo + strips) {   // a whole low or high row
        const bool hi = r >= lo;
        if (hi) r -= lo + strips;
        const uint32_t jj = udiv(r, sg.G, sg.r_g);
        j = hi ? sg.ny - sg.my + jj : jj;
        g = r - jj * sg.G;
    } else {   // an x strip of a box row
        r -= lo;
        const uint32_t jj = udiv(r, sg.gx, sg.r_gx), gg = r - jj * sg.gx;
        j = sg.my + jj;
        g = gg < sg.gi0 ? gg : gg + sg.ngi;
    }
}

// X-strip entry t of a box (groups [0, gi0) and [gi0+ngi, G) of the nj box rows of the nk box
// planes, row after row): every such group lies outside the y and z slabs and off the high
// boundary rows and planes.
__device__ __forceinline__ void strip_pos(const ShellGeom &sg, uint32_t t, uint32_t &g, uint32_t &j, uint32_t &k) {
    const uint32_t r = udiv(t, sg.gx, sg.r_gx), s = t - r * sg.gx;
    const uint32_t kk = udiv(r, sg.nj, sg.r_nj);
    j = sg.my + (r - kk * sg.nj);
    k = sg.mz + kk;
    g = s < sg.gi0 ? s : s + sg.ngi;
}

// Rest row r of the shell, in memory order: the mz low z planes (ny rows each), the my low
// and my high rows of each box plane, the mz high z planes.
__device__ __forceinline__ void rest_row(const ShellGeom &sg, uint32_t r, uint32_t &j, uint32_t &k) {
    if (r < sg.nrlo) {
        k = udiv(r, sg.ny, sg.r_ny);
        j = r - k * sg.ny;
    } else if (r < sg.nrlo + sg.nrmid) {
        const uint32_t u = r - sg.nrlo, m2 = 2 * sg.my;
        const uint32_t kk = udiv(u, m2, sg.r_2my), jj = u - kk * m2;
        k = sg.mz + kk;
        j = jj < sg.my ? jj : sg.ny - m2 + jj;
    } else {
        const uint32_t u = r - sg.nrlo - sg.nrmid, kk = udiv(u, sg.ny, sg.r_ny);
        j = u - kk * sg.ny;
        k = sg.nz - sg.mz + kk;
    }
}

// The group a lane of a vector kernel updates; in_range is false for lanes past the end.
struct LaneGroup { int64_t i0, j, k; bool in_range; };

// Shell modes of the vector kernels: 0 the shell layout (or the plane mapping), 1 the x strips
// of the box rows (no y or z slab), 2 the box's x range of the rest rows (no x slab, no x
// boundary), 3 the x edges of the rest rows.
enum { SH_ALL = 0, SH_STRIP = 1, SH_NOX = 2, SH_EDGE = 3 };
template <int MODE>
__device__ __forceinline__ LaneGroup mode_group(const ShellGeom &sg, int tyv) {
    const uint32_t t = blockIdx.x * (uint32_t)(64 * tyv) + threadIdx.y * 64 + threadIdx.x;
    const uint32_t n = MODE == SH_STRIP ? sg.nstrip : (MODE == SH_NOX ? sg.nnox : sg.nedge);
    LaneGroup r;
    r.in_range = t < n;
    const uint32_t tc = r.in_range ? t : 0u;
    uint32_t g, j, k;
    if (MODE == SH_STRIP) {
        strip_pos(sg, tc, g, j, k);
    } else {
        const uint32_t w = MODE == SH_EDGE ? sg.gx : sg.ngi, rw = MODE == SH_EDGE ? sg.r_gx : sg.r_ngi;
        const uint32_t row = udiv(tc, w, rw), s = tc - row * w;
        g = MODE == SH_EDGE ? (s < sg.gi0 ? s : s + sg.ngi) : sg.gi0 + s;
        rest_row(sg, row, j, k);
    }
    r.i0 = (int64_t)g * VX;
    r.j = (int64_t)j;
    r.k = (int64_t)k;
    return r;
}

// The group of a lane in mode 0: either the flattened plane mapping above (the whole
// volume, one z-plane per workgroup) or shell entry t.  Lanes past the end still load from
// valid addresses and take part in the shuffles.
__device__ __forceinline__ LaneGroup lane_group(const ShellGeom &sg, int64_t nx, int64_t ny, int64_t bpp, int tyv) {
    LaneGroup r;
    if (sg.n) {
        const uint32_t t = blockIdx.x * (uint32_t)(64 * tyv) + threadIdx.y * 64 + threadIdx.x;
        r.in_range = t < sg.n;
        uint32_t g, j, k;
        shell_pos(sg, r.in_range ? t : 0u, g, j, k);
        r.i0 = (int64_t)g * VX;
        r.j = (int64_t)j;
        r.k = (int64_t)k;
    } else {
        const PlaneGroup pg = plane_group(nx, ny, bpp, tyv);
        r.i0 = pg.i0;
        r.j = pg.j;
        r.k = pg.bz;
        r.in_range = pg.in_plane;
    }
    return r;
}

// Whether the neighbouring lane (delta = -1: previous, +1: next) loaded the adjacent group
// of the same x row; only then may its values stand in for a memory read.
__device__ __forceinline__ bool lane_holds(int64_t o, int delta) {
    const uint32_t lo = (uint32_t)o, hi = (uint32_t)((uint64_t)o >> 32);
    const uint32_t nlo = delta < 0 ? __shfl_up(lo, 1) : __shfl_down(lo, 1);
    const uint32_t nhi = delta < 0 ? __shfl_up(hi, 1) : __shfl_down(hi, 1);
    const int lane = threadIdx.x;
    const bool edge = delta < 0 ? lane == 0 : lane == 63;
    return !edge && (int64_t)(((uint64_t)nhi << 32) | nlo) == o + delta * VX;
}

// PP: the updated H goes to Hxo/Hyo/Hzo (the other buffer set of a ping-pong step) instead of
// back to Hx/Hy/Hz; the whole group is stored, so a cell at i = nx-1 carries its old value over.
// MODE (SH_*): SH_STRIP takes the x strips of the box rows, where no y or z slab term exists,
// and SH_NOX the box's x range of the rest rows, where no x slab term exists: those terms and
// their registers drop out, and either kernel fits two waves per SIMD beside the film update,
// where the general form fits one.  SH_EDGE takes the x edges of the rest rows.
template <typename P, int TYV, bool PP, int MODE = SH_ALL>
__global__ __launch_bounds__(64 * TYV) void yee_H_vec_kernel(
    const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
    float *__restrict__ Hx, float *__restrict__ Hy, float *__restrict__ Hz,
    const float *__restrict__ idcx, const float *__restrict__ idcy, const float *__restrict__ idcz,
    CpmlArgs cp, int64_t nx, int64_t ny, int64_t nz, float dt_mu0, int64_t nbx, ShellGeom sg,
    float *__restrict__ Hxo, float *__restrict__ Hyo, float *__restrict__ Hzo) {
    __builtin_amdgcn_s_setprio(3);   // issue ahead of a film update running beside it
    const LaneGroup lg = MODE == SH_ALL ? lane_group(sg, nx, ny, nbx, TYV) : mode_group<MODE>(sg, TYV);
    // the whole wave takes part in the shuffles; out-of-range lanes only idle
    const bool act = lg.in_range && lg.j < ny - 1 && lg.k < nz - 1;
    const int64_t k = act ? lg.k : 0;
    const int64_t sy = nx, sz = nx * ny;
    const float *xk = (const float *)cp.ik[0], *xa = (const float *)cp.a[0], *xb = (const float *)cp.b[0];
    const float *yk = (const float *)cp.ik[1], *ya = (const float *)cp.a[1], *yb = (const float *)cp.b[1];
    const float *zk = (const float *)cp.ik[2], *za = (const float *)cp.a[2], *zb = (const float *)cp.b[2];
    P **ps = (P **)cp.psi;
    const int64_t wx = cp.w[0], wy = cp.w[1], wz = cp.w[2];
    const int64_t ic = act ? lg.i0 : 0;   // clamp addresses of idle lanes
    const int64_t jc = act ? lg.j : (int64_t)0;   // rows 0 and 1 exist (ny >= 2)
    int64_t jsy = 0;
    const int ry = MODE == SH_STRIP ? 0 : pml_slab(jc, ny, wy, &jsy);
    const float4 idx4 = ld4(idcx + ic), xk4 = ld4(xk + ic);
    const float idy_ = idcy[jc], yk_ = yk[jc], ya_ = ya[jc], yb_ = yb[jc];
    const bool xslab = MODE != SH_NOX && wx > 0 && (ic < wx || ic + VX > nx - wx);
    const int64_t o = ic + sy * jc + sz * k;
    const float4 ex_c = ld4(Ex + o), ey_c = ld4(Ey + o);
    const float4 ex_n = ld4(Ex + o + sz), ey_n = ld4(Ey + o + sz);
    const float4 ez_c = ld4(Ez + o);
    const float4 ez_y = ld4(Ez + o + sy), ex_y = ld4(Ex + o + sy);
    float4 hx = ld4nt(Hx + o), hy = ld4nt(Hy + o), hz = ld4nt(Hz + o);
    // x neighbours of element 3: element 0 of the next lane's group, or memory
    float ez_nx = __shfl_down(ez_c.x, 1), ey_nx = __shfl_down(ey_c.x, 1);
    if (!lane_holds(o, 1) && ic + VX < nx) {
        ez_nx = Ez[o + VX];
        ey_nx = Ey[o + VX];
    }
    const float idz_ = idcz[k], zk_ = zk[k];
    int64_t jsz = 0;
    const int rz = MODE == SH_STRIP ? 0 : pml_slab(k, nz, wz, &jsz);
    const float za_ = rz ? za[k] : 0.f, zb_ = rz ? zb[k] : 0.f;
    // y-slab psi of Hx and Hz, z-slab psi of Hx and Hy (lo or hi slab array)
    const int64_t oy0 = ic + nx * (jsy + wy * k), oz0 = ic + nx * (jc + ny * jsz);
    P *const pxy = ps[ry == 2 ? 1 : 0], *const pzy = ps[ry == 2 ? 11 : 10];
    P *const pxz = ps[rz == 2 ? 3 : 2], *const pyz = ps[rz == 2 ? 5 : 4];
    Psi4<P> qxy{}, qzy{}, qxz{}, qyz{};
    if (act && ry) {
        qxy = psi4_load(pxy, oy0);
        qzy = psi4_load(pzy, oy0);
    }
    if (act && rz) {
        qxz = psi4_load(pxz, oz0);
        qyz = psi4_load(pyz, oz0);
    }
    PsiX<P> px;
    psix_load(px, ps, act, xslab, ic, nx, wx, jc + ny * k, xa, xb);
#pragma unroll
    for (int e = 0; e < VX; ++e) {
        const int64_t i = ic + e;
        if (!act || i >= nx - 1) continue;   // H at i = nx-1 keeps its value
        const float ezc = elc(ez_c, e), eyc = elc(ey_c, e), exc = elc(ex_c, e);
        const float ez_xp = e < VX - 1 ? elc(ez_c, e + 1) : ez_nx;
        const float ey_xp = e < VX - 1 ? elc(ey_c, e + 1) : ey_nx;
        const float idx_ = elc(idx4, e), xk_ = elc(xk4, e);
        const float dEz_dy = (elc(ez_y, e) - ezc) * idy_;
        const float dEy_dz = (elc(ey_n, e) - eyc) * idz_;
        const float dEx_dz = (elc(ex_n, e) - exc) * idz_;
        const float dEz_dx = (ez_xp - ezc) * idx_;
        const float dEy_dx = (ey_xp - eyc) * idx_;
        const float dEx_dy = (elc(ex_y, e) - exc) * idy_;
        const int rx = px.r[e];
        const float pHxy = ry ? psi4_step(qxy, e, yb_, ya_, dEz_dy, cp.sc, cp.isc) : 0.f;
        const float pHxz = rz ? psi4_step(qxz, e, zb_, za_, dEy_dz, cp.sc, cp.isc) : 0.f;
        el(hx, e) = el(hx, e) - dt_mu0 * ((dEz_dy * yk_ + pHxy) - (dEy_dz * zk_ + pHxz));
        const float pHyz = rz ? psi4_step(qyz, e, zb_, za_, dEx_dz, cp.sc, cp.isc) : 0.f;
        float pHyx = 0.f, pHzx = 0.f;
        if (rx) {
            pHyx = psix_step(px.q6[e], px.b[e], px.a[e], dEz_dx, cp.sc, cp.isc);
            pHzx = psix_step(px.q8[e], px.b[e], px.a[e], dEy_dx, cp.sc, cp.isc);
        }
        el(hy, e) = el(hy, e) - dt_mu0 * ((dEx_dz * zk_ + pHyz) - (dEz_dx * xk_ + pHyx));
        const float pHzy = ry ? psi4_step(qzy, e, yb_, ya_, dEx_dy, cp.sc, cp.isc) : 0.f;
        el(hz, e) = el(hz, e) - dt_mu0 * ((dEy_dx * xk_ + pHzx) - (dEx_dy * yk_ + pHzy));
    }
    if (act) {
        st4nt((PP ? Hxo : Hx) + o, hx);
        st4nt((PP ? Hyo : Hy) + o, hy);
        st4nt((PP ? Hzo : Hz) + o, hz);
        psix_store(px, ps, xslab);
        if (ry) {
            psi4_store(pxy, oy0, qxy);
            psi4_store(pzy, oy0, qzy);
        }
        if (rz) {
            psi4_store(pxz, oz0, qxz);
            psi4_store(pyz, oz0, qyz);
        }
    }
}

// MI: inv_eps comes from the per-cell byte index and a table of triples staged in LDS
// instead of the three volumes (same values, 1 byte per cell instead of 12).
template <typename P, int TYV, bool MI, bool LIST>
__global__ __launch_bounds__(64 * TYV) void yee_E_vec_kernel(
    float *__restrict__ Ex, float *__restrict__ Ey, float *__restrict__ Ez,
    const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
    float *__restrict__ Dx, float *__restrict__ Dy, float *__restrict__ Dz,
    const float *__restrict__ iex, const float *__restrict__ iey, const float *__restrict__ iez,
    const uint8_t *__restrict__ mat, const float4 *__restrict__ mtab, int nmat,
    const float *__restrict__ iddx, const float *__restrict__ iddy, const float *__restrict__ iddz,
    CpmlArgs cp, int64_t nx, int64_t ny, int64_t nz, int64_t kchunk, float dt, int64_t nbx, ShellGeom sg) {
    __builtin_amdgcn_s_setprio(3);   // issue ahead of a film update running beside it
    __shared__ float4 tab[MI ? 256 : 1];
    if (MI) {
        for (int t = threadIdx.y * 64 + threadIdx.x; t < nmat; t += 64 * TYV) tab[t] = mtab[t];
        __syncthreads();
    }
    const int lane = threadIdx.x;
    int64_t i0, j, k0, k1;
    bool act;
    if (LIST) {
        // one listed group per lane, one plane; every lane stays for the shuffles
        const LaneGroup lg = lane_group(sg, nx, ny, 1, TYV);
        act = lg.in_range && lg.j >= 1 && lg.k >= 1;
        i0 = lg.i0;
        j = lg.j;
        k0 = act ? lg.k : 1;
        k1 = k0 + 1;
    } else {
        const PlaneGroup pg = plane_group(nx, ny, nbx, TYV);
        i0 = pg.i0;
        j = pg.j;
        act = pg.in_plane && j >= 1;
        k0 = max(pg.bz * kchunk, (int64_t)1);
        k1 = min(pg.bz * kchunk + kchunk, nz);
        if (k0 >= k1) return;   // uniform across the workgroup
    }
    const int64_t sy = nx, sz = nx * ny;
    const float *xk = (const float *)cp.ik[0], *xa = (const float *)cp.a[0], *xb = (const float *)cp.b[0];
    const float *yk = (const float *)cp.ik[1], *ya = (const float *)cp.a[1], *yb = (const float *)cp.b[1];
    const float *zk = (const float *)cp.ik[2], *za = (const float *)cp.a[2], *zb = (const float *)cp.b[2];
    P **ps = (P **)cp.psi;
    const int64_t wx = cp.w[0], wy = cp.w[1], wz = cp.w[2];
    const int64_t ic = act ? i0 : 0;
    const int64_t jc = act ? j : (int64_t)1;   // rows 0 and 1 exist (ny >= 2)
    int64_t jsy;
    const int ry = pml_slab(jc, ny, wy, &jsy);
    const float4 idx4 = ld4(iddx + ic), xk4 = ld4(xk + ic);
    const float idy_ = iddy[jc], yk_ = yk[jc], ya_ = ya[jc], yb_ = yb[jc];
    const bool xslab = wx > 0 && (ic < wx || ic + VX > nx - wx);
    int64_t o = ic + sy * jc + sz * k0;
    float4 hx_p = ld4(Hx + o - sz), hy_p = ld4(Hy + o - sz);
    for (int64_t k = k0; k < k1; ++k, o += sz) {
        const float4 hx_c = ld4(Hx + o), hy_c = ld4(Hy + o), hz_c = ld4(Hz + o);
        const float4 hz_y = ld4(Hz + o - sy), hx_y = ld4(Hx + o - sy);
        float4 dx = ld4nt(Dx + o), dy = ld4nt(Dy + o), dz = ld4nt(Dz + o);
        // inv_eps (or the material bytes) load with the fields; the table lookup follows the
        // cell loop, so no load waits before the last one is issued
        float4 iex4, iey4, iez4;
        uint32_t m4 = 0u;
        if (MI) {
            m4 = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(mat + o));
        } else {
            iex4 = ld4nt(iex + o);
            iey4 = ld4nt(iey + o);
            iez4 = ld4nt(iez + o);
        }
        // x neighbours of element 0: element 3 of the previous lane's group, or memory
        float hz_px = __shfl_up(hz_c.w, 1), hy_px = __shfl_up(hy_c.w, 1);
        if ((LIST ? !lane_holds(o, -1) : lane == 0) && ic > 0) {
            hz_px = Hz[o - 1];
            hy_px = Hy[o - 1];
        }
        const float idz_ = iddz[k], zk_ = zk[k];
        int64_t jsz;
        const int rz = pml_slab(k, nz, wz, &jsz);
        const float za_ = rz ? za[k] : 0.f, zb_ = rz ? zb[k] : 0.f;
        // y-slab psi of Dx and Dz, z-slab psi of Dx and Dy (lo or hi slab array)
        const int64_t oy0 = ic + nx * (jsy + wy * k), oz0 = ic + nx * (jc + ny * jsz);
        P *const pxy = ps[ry == 2 ? 1 : 0], *const pzy = ps[ry == 2 ? 11 : 10];
        P *const pxz = ps[rz == 2 ? 3 : 2], *const pyz = ps[rz == 2 ? 5 : 4];
        Psi4<P> qxy{}, qzy{}, qxz{}, qyz{};
        if (act && ry) {
            qxy = psi4_load(pxy, oy0);
            qzy = psi4_load(pzy, oy0);
        }
        if (act && rz) {
            qxz = psi4_load(pxz, oz0);
            qyz = psi4_load(pyz, oz0);
        }
        PsiX<P> px;
        psix_load(px, ps, act, xslab, ic, nx, wx, jc + ny * k, xa, xb);
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            const int64_t i = ic + e;
            if (!act || i < 1) continue;   // E, D at i = 0 keep their values
            const float hxc = elc(hx_c, e), hyc = elc(hy_c, e), hzc = elc(hz_c, e);
            const float hz_xm = e > 0 ? elc(hz_c, e - 1) : hz_px;
            const float hy_xm = e > 0 ? elc(hy_c, e - 1) : hy_px;
            const float idx_ = elc(idx4, e), xk_ = elc(xk4, e);
            const float dHz_dy = (hzc - elc(hz_y, e)) * idy_;
            const float dHy_dz = (hyc - elc(hy_p, e)) * idz_;
            const float dHx_dz = (hxc - elc(hx_p, e)) * idz_;
            const float dHz_dx = (hzc - hz_xm) * idx_;
            const float dHy_dx = (hyc - hy_xm) * idx_;
            const float dHx_dy = (hxc - elc(hx_y, e)) * idy_;
            const int rx = px.r[e];
            const float pDxy = ry ? psi4_step(qxy, e, yb_, ya_, dHz_dy, cp.sc, cp.isc) : 0.f;
            const float pDxz = rz ? psi4_step(qxz, e, zb_, za_, dHy_dz, cp.sc, cp.isc) : 0.f;
            const float cx_term = (dHz_dy * yk_ + pDxy) - (dHy_dz * zk_ + pDxz);
            const float pDyz = rz ? psi4_step(qyz, e, zb_, za_, dHx_dz, cp.sc, cp.isc) : 0.f;
            float pDyx = 0.f, pDzx = 0.f;
            if (rx) {
                pDyx = psix_step(px.q6[e], px.b[e], px.a[e], dHz_dx, cp.sc, cp.isc);
                pDzx = psix_step(px.q8[e], px.b[e], px.a[e], dHy_dx, cp.sc, cp.isc);
            }
            const float cy_term = (dHx_dz * zk_ + pDyz) - (dHz_dx * xk_ + pDyx);
            const float pDzy = ry ? psi4_step(qzy, e, yb_, ya_, dHx_dy, cp.sc, cp.isc) : 0.f;
            const float cz_term = (dHy_dx * xk_ + pDzx) - (dHx_dy * yk_ + pDzy);
            const float dxn = el(dx, e) + dt * cx_term;
            const float dyn = el(dy, e) + dt * cy_term;
            const float dzn = el(dz, e) + dt * cz_term;
            el(dx, e) = dxn;
            el(dy, e) = dyn;
            el(dz, e) = dzn;
        }
        if (MI) {
            const float4 t0 = tab[m4 & 0xffu], t1 = tab[(m4 >> 8) & 0xffu];
            const float4 t2 = tab[(m4 >> 16) & 0xffu], t3 = tab[m4 >> 24];
            iex4 = make_float4(t0.x, t1.x, t2.x, t3.x);
            iey4 = make_float4(t0.y, t1.y, t2.y, t3.y);
            iez4 = make_float4(t0.z, t1.z, t2.z, t3.z);
        }
        // E = D * inv_eps, formed after the loop so E does not occupy registers through it
        if (act && ic > 0) {
            st4nt(Dx + o, dx); st4nt(Dy + o, dy); st4nt(Dz + o, dz);
            st4nt(Ex + o, make_float4(dx.x * iex4.x, dx.y * iex4.y, dx.z * iex4.z, dx.w * iex4.w));
            st4nt(Ey + o, make_float4(dy.x * iey4.x, dy.y * iey4.y, dy.z * iey4.z, dy.w * iey4.w));
            st4nt(Ez + o, make_float4(dz.x * iez4.x, dz.y * iez4.y, dz.z * iez4.z, dz.w * iez4.w));
        } else if (act) {   // the group holding i = 0: that cell is not part of the E update
            for (int e = 1; e < VX; ++e) {
                Dx[o + e] = elc(dx, e); Dy[o + e] = elc(dy, e); Dz[o + e] = elc(dz, e);
                Ex[o + e] = elc(dx, e) * elc(iex4, e);
                Ey[o + e] = elc(dy, e) * elc(iey4, e);
                Ez[o + e] = elc(dz, e) * elc(iez4, e);
            }
        }
        psix_store(px, ps, act && xslab);
        if (act && ry) {
            psi4_store(pxy, oy0, qxy);
            psi4_store(pzy, oy0, qzy);
        }
        if (act && rz) {
            psi4_store(pxz, oz0, qxz);
            psi4_store(pyz, oz0, qyz);
        }
        hx_p = hx_c;
        hy_p = hy_c;
    }
}

// ---- interior kernels ---------------------------------------------------------------------
// The interior box holds the whole float4 groups that lie in no CPML slab and contain no
// boundary cell on any axis, with its x edges on whole 128-byte lines.  There every psi term
// is +0 and no cell is skipped, so these kernels run the reference's arithmetic with the psi
// terms added as +0 (bit for bit what the general kernels compute there) without the slab
// tests, psi offsets and skip branches, in far fewer registers.  Lanes take consecutive
// groups of the box, row after row and plane after plane, so a wave's loads are contiguous
// runs.  The CPML shell around the box is updated by the general kernels, whose lanes decode
// their shell group from the shell's layout; no cache line is written by both launches.
__device__ __forceinline__ void box_pos(uint32_t t, uint32_t per_plane, uint32_t ngi, uint32_t &kk, uint32_t &jj,
                                        uint32_t &gg) {
    kk = t / per_plane;
    const uint32_t r = t - kk * per_plane;
    jj = r / ngi;
    gg = r - jj * ngi;
}

// Loads and stores at a 32-bit byte offset from a uniform base: every volume of the interior
// and fused kernels is then addressed by one offset register and its base in scalar
// registers (the interior E kernel in 63 VGPRs instead of 66: three of its waves fit each
// SIMD beside the film update instead of two; the interior H kernel, at 64 either way, keeps
// 64-bit element offsets, measured 2% faster alone).  The launchers take these kernels only
// for volumes below 4 GiB.
template <typename T>
__device__ __forceinline__ const T &at(const float *base, uint32_t boff) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}
__device__ __forceinline__ float4 ldb4(const float *base, uint32_t boff) { return at<float4>(base, boff); }
__device__ __forceinline__ float4 ldb4nt(const float *base, uint32_t boff) {
    return ld4nt(reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + boff));
}
__device__ __forceinline__ void stb4nt(float *base, uint32_t boff, float4 v) {
    st4nt(reinterpret_cast<float *>(reinterpret_cast<char *>(base) + boff), v);
}

__global__ __launch_bounds__(256) void yee_H_int_kernel(
    const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
    float *__restrict__ Hx, float *__restrict__ Hy, float *__restrict__ Hz,
    const float *__restrict__ idcx, const float *__restrict__ idcy, const float *__restrict__ idcz,
    const float *__restrict__ xk, const float *__restrict__ yk, const float *__restrict__ zk, int64_t nx,
    int64_t ny, uint32_t gi0, uint32_t ngi, uint32_t j0, uint32_t per_plane, uint32_t k0, uint32_t ntot,
    float dt_mu0) {
    __builtin_amdgcn_s_setprio(3);   // issue ahead of a film update running beside it
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const bool act = t < ntot;
    uint32_t kk, jj, gg;
    box_pos(act ? t : 0u, per_plane, ngi, kk, jj, gg);
    const int64_t i0 = (int64_t)(gi0 + gg) * VX, j = (int64_t)(j0 + jj), k = (int64_t)(k0 + kk);
    const int64_t sy = nx, sz = nx * ny;
    const int64_t o = i0 + sy * j + sz * k;
    const float4 ex_c = ld4(Ex + o), ey_c = ld4(Ey + o), ez_c = ld4(Ez + o);
    const float4 ex_n = ld4(Ex + o + sz), ey_n = ld4(Ey + o + sz);
    const float4 ez_y = ld4(Ez + o + sy), ex_y = ld4(Ex + o + sy);
    float4 hx = ld4nt(Hx + o), hy = ld4nt(Hy + o), hz = ld4nt(Hz + o);
    // the next lane holds the next group of this row unless this is the row's last group
    float ez_nx = __shfl_down(ez_c.x, 1), ey_nx = __shfl_down(ey_c.x, 1);
    if ((threadIdx.x & 63) == 63 || gg + 1 == ngi) {
        ez_nx = Ez[o + VX];
        ey_nx = Ey[o + VX];
    }
    const float4 idx4 = ld4(idcx + i0), xk4 = ld4(xk + i0);
    const float idy_ = idcy[j], yk_ = yk[j], idz_ = idcz[k], zk_ = zk[k];
#pragma unroll
    for (int e = 0; e < VX; ++e) {
        const float ezc = elc(ez_c, e), eyc = elc(ey_c, e), exc = elc(ex_c, e);
        const float ez_xp = e < VX - 1 ? elc(ez_c, e + 1) : ez_nx;
        const float ey_xp = e < VX - 1 ? elc(ey_c, e + 1) : ey_nx;
        const float idx_ = elc(idx4, e), xk_ = elc(xk4, e);
        const float dEz_dy = (elc(ez_y, e) - ezc) * idy_;
        const float dEy_dz = (elc(ey_n, e) - eyc) * idz_;
        const float dEx_dz = (elc(ex_n, e) - exc) * idz_;
        const float dEz_dx = (ez_xp - ezc) * idx_;
        const float dEy_dx = (ey_xp - eyc) * idx_;
        const float dEx_dy = (elc(ex_y, e) - exc) * idy_;
        el(hx, e) = el(hx, e) - dt_mu0 * ((dEz_dy * yk_ + 0.0f) - (dEy_dz * zk_ + 0.0f));
        el(hy, e) = el(hy, e) - dt_mu0 * ((dEx_dz * zk_ + 0.0f) - (dEz_dx * xk_ + 0.0f));
        el(hz, e) = el(hz, e) - dt_mu0 * ((dEy_dx * xk_ + 0.0f) - (dEx_dy * yk_ + 0.0f));
    }
    if (act) {
        st4nt(Hx + o, hx);
        st4nt(Hy + o, hy);
        st4nt(Hz + o, hz);
    }
}

template <bool MI>
__global__ __launch_bounds__(256) void yee_E_int_kernel(
    float *__restrict__ Ex, float *__restrict__ Ey, float *__restrict__ Ez,
    const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
    float *__restrict__ Dx, float *__restrict__ Dy, float *__restrict__ Dz,
    const float *__restrict__ iex, const float *__restrict__ iey, const float *__restrict__ iez,
    const uint8_t *__restrict__ mat, const float4 *__restrict__ mtab, int nmat,
    const float *__restrict__ iddx, const float *__restrict__ iddy, const float *__restrict__ iddz,
    const float *__restrict__ xk, const float *__restrict__ yk, const float *__restrict__ zk, int64_t nx,
    int64_t ny, uint32_t gi0, uint32_t ngi, uint32_t j0, uint32_t per_plane, uint32_t k0, uint32_t ntot,
    float dt) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ float4 tab[MI ? 256 : 1];
    if (MI) {
        for (int t = threadIdx.x; t < nmat; t += 256) tab[t] = mtab[t];
        __syncthreads();
    }
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const bool act = t < ntot;
    uint32_t kk, jj, gg;
    box_pos(act ? t : 0u, per_plane, ngi, kk, jj, gg);
    const uint32_t i0 = (gi0 + gg) * VX, j = j0 + jj, k = k0 + kk;
    const uint32_t sy = (uint32_t)nx, sz = (uint32_t)(nx * ny);
    const uint32_t ob = (i0 + sy * j + sz * k) * 4u;
    const float4 hx_p = ldb4(Hx - sz, ob), hy_p = ldb4(Hy - sz, ob);
    const float4 hx_c = ldb4(Hx, ob), hy_c = ldb4(Hy, ob), hz_c = ldb4(Hz, ob);
    const float4 hz_y = ldb4(Hz - sy, ob), hx_y = ldb4(Hx - sy, ob);
    float4 dx = ldb4nt(Dx, ob), dy = ldb4nt(Dy, ob), dz = ldb4nt(Dz, ob);
    // the previous lane holds the previous group of this row unless this is its first group
    float hz_px = __shfl_up(hz_c.w, 1), hy_px = __shfl_up(hy_c.w, 1);
    if ((threadIdx.x & 63) == 0 || gg == 0) {
        hz_px = at<float>(Hz - 1, ob);
        hy_px = at<float>(Hy - 1, ob);
    }
    const float4 idx4 = ldb4(iddx, i0 * 4u), xk4 = ldb4(xk, i0 * 4u);
    const float idy_ = iddy[j], yk_ = yk[j], idz_ = iddz[k], zk_ = zk[k];
    // inv_eps (or the material bytes) load with the fields; the table lookup follows the cell
    // loop, so no load waits before the last one is issued
    float4 iex4, iey4, iez4;
    uint32_t m4 = 0u;
    if (MI) {
        m4 = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(mat + (ob >> 2)));
    } else {
        iex4 = ldb4nt(iex, ob);
        iey4 = ldb4nt(iey, ob);
        iez4 = ldb4nt(iez, ob);
    }
#pragma unroll
    for (int e = 0; e < VX; ++e) {
        const float hxc = elc(hx_c, e), hyc = elc(hy_c, e), hzc = elc(hz_c, e);
        const float hz_xm = e > 0 ? elc(hz_c, e - 1) : hz_px;
        const float hy_xm = e > 0 ? elc(hy_c, e - 1) : hy_px;
        const float idx_ = elc(idx4, e), xk_ = elc(xk4, e);
        const float dHz_dy = (hzc - elc(hz_y, e)) * idy_;
        const float dHy_dz = (hyc - elc(hy_p, e)) * idz_;
        const float dHx_dz = (hxc - elc(hx_p, e)) * idz_;
        const float dHz_dx = (hzc - hz_xm) * idx_;
        const float dHy_dx = (hyc - hy_xm) * idx_;
        const float dHx_dy = (hxc - elc(hx_y, e)) * idy_;
        const float cx_term = (dHz_dy * yk_ + 0.0f) - (dHy_dz * zk_ + 0.0f);
        const float cy_term = (dHx_dz * zk_ + 0.0f) - (dHz_dx * xk_ + 0.0f);
        const float cz_term = (dHy_dx * xk_ + 0.0f) - (dHx_dy * yk_ + 0.0f);
        el(dx, e) = el(dx, e) + dt * cx_term;
        el(dy, e) = el(dy, e) + dt * cy_term;
        el(dz, e) = el(dz, e) + dt * cz_term;
    }
    if (MI) {
        const float4 t0 = tab[m4 & 0xffu], t1 = tab[(m4 >> 8) & 0xffu];
        const float4 t2 = tab[(m4 >> 16) & 0xffu], t3 = tab[m4 >> 24];
        iex4 = make_float4(t0.x, t1.x, t2.x, t3.x);
        iey4 = make_float4(t0.y, t1.y, t2.y, t3.y);
        iez4 = make_float4(t0.z, t1.z, t2.z, t3.z);
    }
    if (act) {
        stb4nt(Dx, ob, dx); stb4nt(Dy, ob, dy); stb4nt(Dz, ob, dz);
        stb4nt(Ex, ob, make_float4(dx.x * iex4.x, dx.y * iex4.y, dx.z * iex4.z, dx.w * iex4.w));
        stb4nt(Ey, ob, make_float4(dy.x * iey4.x, dy.y * iey4.y, dy.z * iey4.z, dy.w * iey4.w));
        stb4nt(Ez, ob, make_float4(dz.x * iez4.x, dz.y * iez4.y, dz.z * iez4.z, dz.w * iez4.w));
    }
}

// The six E and H volumes of one buffer set (ex, ey, ez, hx, hy, hz).
struct PlanePtrs {
    float *p[6];
};

// ---- fused interior H + E (ping-pong) -----------------------------------------------------
// One pass over the interior box for a whole step: H^{n+1/2} and then E^{n+1}, reading E^n and
// H^{n-1/2} from one buffer set and writing H^{n+1/2}, E^{n+1} to the other (D in place), so
// no value a workgroup reads is written during the launch.  A workgroup owns T = FL - ngi
// consecutive float4 groups of a box plane (rows flattened, row after row) and marches them
// through a chunk of planes.  Its lanes 0..ngi-1 take the ngi groups just before (the row
// below each of its first ngi groups): they recompute that row's H, which the E update needs
// as its y neighbour, from the same unmodified inputs, so the values equal the owner's bit for
// bit.  Each plane: E^n of the plane is staged in LDS, every lane forms its group's H from its
// own and its neighbours' E (x and y neighbours from LDS, z neighbour from registers), the new
// H is staged in LDS, and the owning lanes form D and E from it (x and y neighbours from LDS,
// the z neighbour kept from the previous plane).  Neighbours outside the box are CPML-shell
// cells: their new H was written by the shell launch that runs first, and their E^n is read
// from the input set.  A chunk that starts inside the box first recomputes the new H of the
// plane below it.  Per cell this reads E, H, D and the material byte once and writes H, D and
// E once: 73 bytes instead of the 85 of the two interior launches.  The per-cell arithmetic is
// that of yee_H_int_kernel and yee_E_int_kernel, operation for operation.
constexpr int FL = 512;   // lanes per workgroup (8 waves: two fit each SIMD beside the film update)

template <bool MI>
__global__ __launch_bounds__(FL) void yee_HE_int_kernel(
    const float *__restrict__ Ex, const float *__restrict__ Ey, const float *__restrict__ Ez,
    const float *__restrict__ Hx, const float *__restrict__ Hy, const float *__restrict__ Hz,
    float *__restrict__ Exo, float *__restrict__ Eyo, float *__restrict__ Ezo,
    float *__restrict__ Hxo, float *__restrict__ Hyo, float *__restrict__ Hzo,
    float *__restrict__ Dx, float *__restrict__ Dy, float *__restrict__ Dz,
    const float *__restrict__ iex, const float *__restrict__ iey, const float *__restrict__ iez,
    const uint8_t *__restrict__ mat, const float4 *__restrict__ mtab, int nmat,
    const float *__restrict__ idcx, const float *__restrict__ idcy, const float *__restrict__ idcz,
    const float *__restrict__ iddx, const float *__restrict__ iddy, const float *__restrict__ iddz,
    const float *__restrict__ xk, const float *__restrict__ yk, const float *__restrict__ zk, uint32_t nx,
    uint32_t ny, uint32_t gi0, uint32_t ngi, uint32_t j0, uint32_t nplane, uint32_t k0, uint32_t kend,
    uint32_t tiles, uint32_t kchunk, float dt_mu0, float dt) {
    __builtin_amdgcn_s_setprio(3);   // issue ahead of a film update running beside it
    __shared__ float4 sE[3][FL];
    __shared__ float4 sH[3][FL];
    __shared__ float4 tab[MI ? 256 : 1];
    const uint32_t l = threadIdx.x;
    if (MI)   // read after the first barrier of the plane loop
        for (int t = (int)l; t < nmat; t += FL) tab[t] = mtab[t];
    const uint32_t tile = blockIdx.x % tiles, chunk = blockIdx.x / tiles;
    const int64_t f = (int64_t)tile * (FL - ngi) - (int64_t)ngi + (int64_t)l;   // group of the box plane
    const bool valid = f >= 0 && f < (int64_t)nplane;
    const bool own = valid && l >= ngi;
    const uint32_t fc = valid ? (uint32_t)f : 0u;
    const uint32_t jj = fc / ngi, gg = fc - jj * ngi;
    const uint32_t i0 = (gi0 + gg) * VX, j = j0 + jj;
    // byte offsets: x step 4, row step sy, plane step sz
    const uint32_t sy = 4u * nx, sz = sy * ny;
    const uint32_t brow = 4u * i0 + sy * j;
    const uint32_t kc0 = k0 + chunk * kchunk, kc1 = min(kc0 + kchunk, kend);
    const float idcy_ = idcy[j], iddy_ = iddy[j], yk_ = yk[j];
    // E^n of the current plane and the new H of the plane below (own groups)
    uint32_t k = kc0 == k0 ? kc0 : kc0 - 1;
    float4 ecx = make_float4(0.f, 0.f, 0.f, 0.f), ecy = ecx, ecz = ecx, hpx = ecx, hpy = ecx;
    if (valid) {
        const uint32_t o = brow + sz * k;
        ecx = ldb4(Ex, o); ecy = ldb4(Ey, o); ecz = ldb4(Ez, o);
    }
    if (kc0 == k0 && own) {   // the plane below the box is CPML shell, updated by the shell launch
        hpx = ldb4(Hxo, brow + sz * (k0 - 1));
        hpy = ldb4(Hyo, brow + sz * (k0 - 1));
    }
    for (; k < kc1; ++k) {
        const bool upd = k >= kc0;   // false on the recomputed plane below the chunk
        const uint32_t o = brow + sz * k;
        float4 enx = ecx, eny = ecy, enz = ecz, hx = ecx, hy = ecy, hz = ecz;
        if (valid) {
            enx = ldb4(Ex, o + sz); eny = ldb4(Ey, o + sz); enz = ldb4(Ez, o + sz);
            hx = ldb4nt(Hx, o); hy = ldb4nt(Hy, o); hz = ldb4nt(Hz, o);
        }
        sE[0][l] = ecx; sE[1][l] = ecy; sE[2][l] = ecz;
        __syncthreads();
        // D and the material of the plane, for the E half: issued here, after the barrier, so
        // they are in flight during the H arithmetic without holding registers across it
        float4 dx = ecx, dy = ecy, dz = ecz, iex4 = ecx, iey4 = ecy, iez4 = ecz;
        uint32_t m4 = 0u;
        if (upd && own) {
            dx = ldb4nt(Dx, o); dy = ldb4nt(Dy, o); dz = ldb4nt(Dz, o);
            if (MI) {
                m4 = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(mat + (o >> 2)));
            } else {
                iex4 = ldb4nt(iex, o); iey4 = ldb4nt(iey, o); iez4 = ldb4nt(iez, o);
            }
        }
        if (valid) {
            // y neighbour (next row): the lane ngi on, or memory past the workgroup / the box
            float4 ez_y, ex_y;
            if (l + ngi < (uint32_t)FL && f + ngi < (int64_t)nplane) {
                ez_y = sE[2][l + ngi];
                ex_y = sE[0][l + ngi];
            } else {
                ez_y = ldb4(Ez, o + sy);
                ex_y = ldb4(Ex, o + sy);
            }
            // x neighbour of element 3: the next lane's group in this row, or memory
            float ez_nx, ey_nx;
            if (gg + 1 < ngi && l + 1 < (uint32_t)FL) {
                ez_nx = sE[2][l + 1].x;
                ey_nx = sE[1][l + 1].x;
            } else {
                ez_nx = at<float>(Ez, o + 4u * VX);
                ey_nx = at<float>(Ey, o + 4u * VX);
            }
            const float idz_ = idcz[k], zk_ = zk[k];
            const float4 idc4 = ldb4(idcx, 4u * i0), xk4 = ldb4(xk, 4u * i0);
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const float ezc = elc(ecz, e), eyc = elc(ecy, e), exc = elc(ecx, e);
                const float ez_xp = e < VX - 1 ? elc(ecz, e + 1) : ez_nx;
                const float ey_xp = e < VX - 1 ? elc(ecy, e + 1) : ey_nx;
                const float idx_ = elc(idc4, e), xk_ = elc(xk4, e);
                const float dEz_dy = (elc(ez_y, e) - ezc) * idcy_;
                const float dEy_dz = (elc(eny, e) - eyc) * idz_;
                const float dEx_dz = (elc(enx, e) - exc) * idz_;
                const float dEz_dx = (ez_xp - ezc) * idx_;
                const float dEy_dx = (ey_xp - eyc) * idx_;
                const float dEx_dy = (elc(ex_y, e) - exc) * idcy_;
                el(hx, e) = el(hx, e) - dt_mu0 * ((dEz_dy * yk_ + 0.0f) - (dEy_dz * zk_ + 0.0f));
                el(hy, e) = el(hy, e) - dt_mu0 * ((dEx_dz * zk_ + 0.0f) - (dEz_dx * xk_ + 0.0f));
                el(hz, e) = el(hz, e) - dt_mu0 * ((dEy_dx * xk_ + 0.0f) - (dEx_dy * yk_ + 0.0f));
            }
        }
        sH[0][l] = hx; sH[1][l] = hy; sH[2][l] = hz;
        __syncthreads();
        if (upd && own) {
            // y neighbour (previous row): the lane ngi back (the recomputed row for the first
            // ngi groups), or the shell row below the box
            float4 hz_y, hx_y;
            if (jj > 0) {
                hz_y = sH[2][l - ngi];
                hx_y = sH[0][l - ngi];
            } else {
                hz_y = ldb4(Hzo, o - sy);
                hx_y = ldb4(Hxo, o - sy);
            }
            // x neighbour of element 0: the previous lane's group in this row, or the shell
            float hz_px, hy_px;
            if (gg > 0) {
                hz_px = sH[2][l - 1].w;
                hy_px = sH[1][l - 1].w;
            } else {
                hz_px = at<float>(Hzo, o - 4u);
                hy_px = at<float>(Hyo, o - 4u);
            }
            const float idz_ = iddz[k], zk_ = zk[k];
            const float4 idd4 = ldb4(iddx, 4u * i0), xk4 = ldb4(xk, 4u * i0);
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const float hxc = elc(hx, e), hyc = elc(hy, e), hzc = elc(hz, e);
                const float hz_xm = e > 0 ? elc(hz, e - 1) : hz_px;
                const float hy_xm = e > 0 ? elc(hy, e - 1) : hy_px;
                const float idx_ = elc(idd4, e), xk_ = elc(xk4, e);
                const float dHz_dy = (hzc - elc(hz_y, e)) * iddy_;
                const float dHy_dz = (hyc - elc(hpy, e)) * idz_;
                const float dHx_dz = (hxc - elc(hpx, e)) * idz_;
                const float dHz_dx = (hzc - hz_xm) * idx_;
                const float dHy_dx = (hyc - hy_xm) * idx_;
                const float dHx_dy = (hxc - elc(hx_y, e)) * iddy_;
                const float cx_term = (dHz_dy * yk_ + 0.0f) - (dHy_dz * zk_ + 0.0f);
                const float cy_term = (dHx_dz * zk_ + 0.0f) - (dHz_dx * xk_ + 0.0f);
                const float cz_term = (dHy_dx * xk_ + 0.0f) - (dHx_dy * yk_ + 0.0f);
                el(dx, e) = el(dx, e) + dt * cx_term;
                el(dy, e) = el(dy, e) + dt * cy_term;
                el(dz, e) = el(dz, e) + dt * cz_term;
            }
            if (MI) {
                const float4 t0 = tab[m4 & 0xffu], t1 = tab[(m4 >> 8) & 0xffu];
                const float4 t2 = tab[(m4 >> 16) & 0xffu], t3 = tab[m4 >> 24];
                iex4 = make_float4(t0.x, t1.x, t2.x, t3.x);
                iey4 = make_float4(t0.y, t1.y, t2.y, t3.y);
                iez4 = make_float4(t0.z, t1.z, t2.z, t3.z);
            }
            stb4nt(Hxo, o, hx); stb4nt(Hyo, o, hy); stb4nt(Hzo, o, hz);
            stb4nt(Dx, o, dx); stb4nt(Dy, o, dy); stb4nt(Dz, o, dz);
            stb4nt(Exo, o, make_float4(dx.x * iex4.x, dx.y * iex4.y, dx.z * iex4.z, dx.w * iex4.w));
            stb4nt(Eyo, o, make_float4(dy.x * iey4.x, dy.y * iey4.y, dy.z * iey4.z, dy.w * iey4.w));
            stb4nt(Ezo, o, make_float4(dz.x * iez4.x, dz.y * iez4.y, dz.z * iez4.z, dz.w * iez4.w));
        }
        hpx = hx; hpy = hy;
        ecx = enx; ecy = eny; ecz = enz;
    }
}

// The cells no update writes — E on the planes i = 0, j = 0, k = 0 and H on the planes
// i = nx-1, j = ny-1, k = nz-1 — copied from one buffer set to the other.  blockIdx.y: 0-2
// the E planes normal to x, y, z; 3-5 the H planes (a grid of 3 copies E only).
__global__ __launch_bounds__(256) void sync_boundary_kernel(PlanePtrs a, PlanePtrs b, int64_t nx, int64_t ny,
                                                            int64_t nz) {
    const int q = blockIdx.y, axis = q % 3;
    const bool h = q >= 3;
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
    if (c >= n1 * n2) return;
    const int64_t u = c % n1, v = c / n1;
    const int64_t idx = h ? (axis == 0 ? nx - 1 : (axis == 1 ? ny - 1 : nz - 1)) : 0;
    int64_t o;
    if (axis == 0) o = idx + nx * (u + ny * v);
    else if (axis == 1) o = u + nx * (idx + ny * v);
    else o = u + nx * (v + ny * idx);
    for (int comp = 0; comp < 3; ++comp) {
        const int s = (h ? 3 : 0) + comp;
        b.p[s][o] = a.p[s][o];
    }
}

// z-chunk length: enough workgroups to fill 256 CUs several times over.
static int64_t pick_kchunk(int64_t nx, int64_t ny, int64_t nz) {
    const int64_t tiles = (int64_t)cdiv(nx, TX) * (int64_t)cdiv(ny, TY);
    const int64_t target = 8192;  // workgroups per launch
    int64_t nchunks = std::max<int64_t>(1, std::min<int64_t>(nz, target / std::max<int64_t>(tiles, 1)));
    int64_t kc = (nz + nchunks - 1) / nchunks;
    return std::max<int64_t>(kc, 4);
}

static CpmlArgs make_cpml_args(const mp_cpml *c, bool h) {
    CpmlArgs a;
    for (int d = 0; d < 3; ++d) {
        a.ik[d] = c->inv_kappa[d];
        a.a[d] = c->a[d];
        a.b[d] = c->b[d];
    }
    for (int s = 0; s < 12; ++s) a.psi[s] = h ? c->psi_h[s] : c->psi_d[s];
    a.w[0] = c->npml_x;
    a.w[1] = c->npml_y;
    a.w[2] = c->npml_z;
    const double sc = (c->psi_dtype == MP_BF16 || c->psi_dtype == MP_F16) ? c->psi_scale : 1.0;
    a.sc = (float)(sc > 0 ? sc : 1.0);
    a.isc = (float)(1.0 / (sc > 0 ? sc : 1.0));
    return a;
}

static int check_yee(const mp_yee *g, const mp_cpml *c, const char *who) {
    if (!g || !c) return fail_arg("%s: null argument", who);
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail_arg("%s: bad dims", who);
    if (g->nx * g->ny * g->nz >= (int64_t)1 << 40) return fail_arg("%s: grid too large", who);
    const int64_t w[3] = {c->npml_x, c->npml_y, c->npml_z};
    const int64_t n[3] = {g->nx, g->ny, g->nz};
    for (int d = 0; d < 3; ++d) {
        if (w[d] < 0 || (w[d] > 0 && n[d] < 2 * w[d] + 1)) return fail_arg("%s: bad CPML width on axis %d", who, d);
        if (!c->inv_kappa[d] || !c->a[d] || !c->b[d]) return fail_arg("%s: missing CPML coefficients", who);
    }
    return MP_OK;
}

// Rows of 64 lanes per workgroup of the vector kernels (2 and 8 measured no faster).
constexpr int TYV = 4;

// Interior box (groups gi0 .. gi0+ngi-1 of rows j0 .. j0+nj-1 of planes k0 .. k0+nk-1): the
// float4 groups outside every CPML slab and off every boundary cell (a margin of at least one
// cell on each face, so i, j, k >= 1 and <= n-2 throughout).
struct Box {
    int64_t gi0, ngi, j0, nj, k0, nk;
    int64_t n() const { return ngi * nj * nk; }
};
static Box interior_box(const mp_yee *g, const mp_cpml *c) {
    const int64_t mx = std::max<int64_t>(c->npml_x, 1), my = std::max<int64_t>(c->npml_y, 1);
    const int64_t mz = std::max<int64_t>(c->npml_z, 1);
    Box b;
    // x edges on whole 128-byte lines (8 groups; rows start on a line when nx is a multiple of
    // 32), so no cache line is written by both launches: line-sharing edges measured 580
    // against 514 us for the H update of the 512 x 512 x 256 step
    b.gi0 = ((mx + VX - 1) / VX + 7) / 8 * 8;
    b.ngi = std::max<int64_t>((g->nx - mx) / VX / 8 * 8 - b.gi0, 0);
    b.j0 = my;
    b.nj = std::max<int64_t>(g->ny - 2 * my, 0);
    b.k0 = mz;
    b.nk = std::max<int64_t>(g->nz - 2 * mz, 0);
    return b;
}

// The H and E updates split into interior and shell launches when the box is non-empty and
// every group index, lanes past the end included, fits 32 bits; kernel_form
// MP_YEE_WHOLE keeps one launch over the whole volume.
// The interior kernels address the volumes by 32-bit byte offsets: volumes of 4 GiB or more
// (2^30 cells) take the general kernels over the whole volume.
static bool split_ok(const mp_yee *g, const Box &b) {
    return g->kernel_form != MP_YEE_WHOLE && b.n() > 0 && 4 * g->nx * g->ny * g->nz < ((int64_t)1 << 32);
}

// Shell layout of a box (see ShellGeom); n = 0 when a divisor would be below 2, where the
// launchers keep the whole-volume launch.
// strips_apart: the x strips of the box rows are left out of the layout (sgx = 0) for a
// launch of their own (nstrip groups).
static ShellGeom shell_geom(const mp_yee *g, const Box &b, bool strips_apart = false) {
    ShellGeom sg{};
    const int64_t G = g->nx / VX, ny = g->ny, nz = g->nz;
    const int64_t pfull = ny * G, gx = G - b.ngi, sgx = strips_apart ? 0 : gx;
    const int64_t pint = 2 * b.j0 * G + b.nj * sgx, nstrip = b.nj * b.nk * gx;
    const int64_t nlo = b.k0 * pfull, nmid = b.nk * pint, n = 2 * nlo + nmid;
    if (pfull < 2 || pint < 2 || G < 2 || gx < 2 || (strips_apart && b.nj < 2) ||
        n + (strips_apart ? nstrip : 0) != G * ny * nz - b.n())
        return sg;
    auto recip = [](int64_t d) { return (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d); };
    sg.n = (uint32_t)n;
    sg.G = (uint32_t)G;
    sg.ny = (uint32_t)ny;
    sg.nz = (uint32_t)nz;
    sg.gi0 = (uint32_t)b.gi0;
    sg.ngi = (uint32_t)b.ngi;
    sg.my = (uint32_t)b.j0;
    sg.nj = (uint32_t)b.nj;
    sg.mz = (uint32_t)b.k0;
    sg.pfull = (uint32_t)pfull;
    sg.pint = (uint32_t)pint;
    sg.gx = (uint32_t)gx;
    sg.nlo = (uint32_t)nlo;
    sg.nmid = (uint32_t)nmid;
    sg.r_pfull = recip(pfull);
    sg.r_pint = recip(pint);
    sg.r_g = recip(G);
    sg.r_gx = recip(gx);
    sg.sgx = (uint32_t)sgx;
    sg.nstrip = (uint32_t)nstrip;
    sg.r_nj = recip(b.nj);
    // the rest rows by x (strips apart only): every divisor at least 2, and the three parts
    // with the box cover the volume
    const int64_t nrlo = b.k0 * ny, nrmid = b.nk * 2 * b.j0, rows = 2 * nrlo + nrmid;
    if (strips_apart && ny >= 2 && b.ngi >= 2 && b.j0 >= 1 && rows * G + nstrip + b.n() == G * ny * nz) {
        sg.nrlo = (uint32_t)nrlo;
        sg.nrmid = (uint32_t)nrmid;
        sg.nnox = (uint32_t)(rows * b.ngi);
        sg.nedge = (uint32_t)(rows * gx);
        sg.r_ny = recip(ny);
        sg.r_2my = recip(2 * b.j0);
        sg.r_ngi = recip(b.ngi);
    }
    return sg;
}

static bool g_split_x = false;
void yee_shell_split_x(bool on) { g_split_x = on; }

// pp (ping-pong step): only the CPML-shell launch, from g's E and H into pp's H; the interior
// is then updated by the fused kernel.
template <typename P, int TY>
static int launch_H_vec_t(const mp_yee *g, const mp_cpml *c, double dt, double mu0, const mp_yee *pp = nullptr) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    if (ny < 2 || nz < 2) return MP_OK;   // empty update range (Kernels.jl:60-101)
    const float dt_mu0 = (float)dt / (float)mu0;
    const Box b = interior_box(g, c);
    const ShellGeom sg = split_ok(g, b) ? shell_geom(g, b) : ShellGeom{};
    dim3 block(64, TY, 1);
    if (pp) {
        if (!sg.n) return fail_arg("fused Yee step: no interior/shell split for this grid");
        KernelTimer t("yee_H");
        yee_H_vec_kernel<P, TY, true><<<cdiv(sg.n, 64 * TY), block, 0, stream()>>>(
            (const float *)g->ex, (const float *)g->ey, (const float *)g->ez, (float *)g->hx, (float *)g->hy,
            (float *)g->hz, (const float *)g->inv_d_cell_x, (const float *)g->inv_d_cell_y,
            (const float *)g->inv_d_cell_z, make_cpml_args(c, true), nx, ny, nz, dt_mu0, 1, sg, (float *)pp->hx,
            (float *)pp->hy, (float *)pp->hz);
        MP_LAUNCH_CHECK("yee_H_vec_kernel");
        return MP_OK;
    }
    if (!sg.n) {
        // one launch: one z-plane per workgroup range (nbx workgroups per plane)
        const int64_t nbx = cdiv((nx / VX) * ny, 64 * TY), nbz = nz;
        KernelTimer t("yee_H");
        yee_H_vec_kernel<P, TY, false><<<(unsigned)(nbx * nbz), block, 0, stream()>>>(
            (const float *)g->ex, (const float *)g->ey, (const float *)g->ez, (float *)g->hx, (float *)g->hy,
            (float *)g->hz, (const float *)g->inv_d_cell_x, (const float *)g->inv_d_cell_y,
            (const float *)g->inv_d_cell_z, make_cpml_args(c, true), nx, ny, nz, dt_mu0, nbx, ShellGeom{},
            nullptr, nullptr, nullptr);
        MP_LAUNCH_CHECK("yee_H_vec_kernel");
        return MP_OK;
    }
    KernelTimer t("yee_H");
    // the x strips of the box rows in a launch of their own, the rest of the shell in another
    // (split by x beside the film update where the layout allows: the box's x range without
    // x-slab terms, then the x edges; 512 x 512 x 256 pump step 1.534 -> 1.518 ms, while alone
    // the split costs the 384^3 stencil 1.045 -> 1.052 ms); strips, interior, rest measured
    // 1.486 ms per 512 x 512 x 256 pump step against 1.488-1.491 for the other orders
    const ShellGeom ss = shell_geom(g, b, true);
    const ShellGeom &sr = ss.n ? ss : sg;
    if (ss.n) {
        yee_H_vec_kernel<P, TY, false, SH_STRIP><<<cdiv(ss.nstrip, 64 * TY), block, 0, stream()>>>(
            (const float *)g->ex, (const float *)g->ey, (const float *)g->ez, (float *)g->hx, (float *)g->hy,
            (float *)g->hz, (const float *)g->inv_d_cell_x, (const float *)g->inv_d_cell_y,
            (const float *)g->inv_d_cell_z, make_cpml_args(c, true), nx, ny, nz, dt_mu0, 1, ss, nullptr, nullptr,
            nullptr);
        MP_LAUNCH_CHECK("yee_H_vec_kernel");
    }
    yee_H_int_kernel<<<cdiv(b.n(), 256), 256, 0, stream()>>>(
        (const float *)g->ex, (const float *)g->ey, (const float *)g->ez, (float *)g->hx, (float *)g->hy,
        (float *)g->hz, (const float *)g->inv_d_cell_x, (const float *)g->inv_d_cell_y,
        (const float *)g->inv_d_cell_z, (const float *)c->inv_kappa[0], (const float *)c->inv_kappa[1],
        (const float *)c->inv_kappa[2], nx, ny, (uint32_t)b.gi0, (uint32_t)b.ngi, (uint32_t)b.j0,
        (uint32_t)(b.ngi * b.nj), (uint32_t)b.k0, (uint32_t)b.n(), dt_mu0);
    MP_LAUNCH_CHECK("yee_H_int_kernel");
#define H_SHELL_ARGS(geom)                                                                                  \
    (const float *)g->ex, (const float *)g->ey, (const float *)g->ez, (float *)g->hx, (float *)g->hy,          \
        (float *)g->hz, (const float *)g->inv_d_cell_x, (const float *)g->inv_d_cell_y,                        \
        (const float *)g->inv_d_cell_z, make_cpml_args(c, true), nx, ny, nz, dt_mu0, 1, geom, nullptr, nullptr, \
        nullptr
    if (g_split_x && ss.n && ss.nnox) {
        yee_H_vec_kernel<P, TY, false, SH_NOX><<<cdiv(ss.nnox, 64 * TY), block, 0, stream()>>>(H_SHELL_ARGS(ss));
        MP_LAUNCH_CHECK("yee_H_vec_kernel");
        yee_H_vec_kernel<P, TY, false, SH_EDGE><<<cdiv(ss.nedge, 64 * TY), block, 0, stream()>>>(H_SHELL_ARGS(ss));
    } else {
        yee_H_vec_kernel<P, TY, false><<<cdiv(sr.n, 64 * TY), block, 0, stream()>>>(H_SHELL_ARGS(sr));
    }
#undef H_SHELL_ARGS
    MP_LAUNCH_CHECK("yee_H_vec_kernel");
    return MP_OK;
}

// shell_only (ping-pong step): only the CPML-shell launch; the fused kernel did the interior.
template <typename P, int TY>
static int launch_E_vec_t(const mp_yee *g, const mp_cpml *c, double dt, bool shell_only = false) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    if (ny < 2 || nz < 2) return MP_OK;   // empty update range (Kernels.jl:103-152)
    const bool mi = g->eps_index && g->eps_table && g->n_eps > 0 && g->n_eps <= 256;
    const Box b = interior_box(g, c);
    const ShellGeom sg = split_ok(g, b) ? shell_geom(g, b) : ShellGeom{};
    dim3 block(64, TY, 1);
#define E_ARGS                                                                                             \
    (float *)g->ex, (float *)g->ey, (float *)g->ez, (const float *)g->hx, (const float *)g->hy,            \
        (const float *)g->hz, (float *)g->dx, (float *)g->dy, (float *)g->dz, (const float *)g->inv_eps_x, \
        (const float *)g->inv_eps_y, (const float *)g->inv_eps_z, g->eps_index,                            \
        (const float4 *)g->eps_table, g->n_eps, (const float *)g->inv_d_dual_x,                            \
        (const float *)g->inv_d_dual_y, (const float *)g->inv_d_dual_z
    if (shell_only && !sg.n) return fail_arg("fused Yee step: no interior/shell split for this grid");
    if (!sg.n) {
        // one launch: one z-plane per workgroup range (nbx workgroups per plane)
        const int64_t nbx = cdiv((nx / VX) * ny, 64 * TY), nbz = nz;
        const unsigned grid = (unsigned)(nbx * nbz);
        KernelTimer t("yee_E");
        if (mi)
            yee_E_vec_kernel<P, TY, true, false><<<grid, block, 0, stream()>>>(
                E_ARGS, make_cpml_args(c, false), nx, ny, nz, 1, (float)dt, nbx, ShellGeom{});
        else
            yee_E_vec_kernel<P, TY, false, false><<<grid, block, 0, stream()>>>(
                E_ARGS, make_cpml_args(c, false), nx, ny, nz, 1, (float)dt, nbx, ShellGeom{});
        MP_LAUNCH_CHECK("yee_E_vec_kernel");
        return MP_OK;
    }
    KernelTimer t("yee_E");
#define INT_ARGS                                                                                          \
    (const float *)c->inv_kappa[0], (const float *)c->inv_kappa[1], (const float *)c->inv_kappa[2], nx, ny, \
        (uint32_t)b.gi0, (uint32_t)b.ngi, (uint32_t)b.j0, (uint32_t)(b.ngi * b.nj), (uint32_t)b.k0,         \
        (uint32_t)b.n(), (float)dt
    if (!shell_only) {
        if (mi) yee_E_int_kernel<true><<<cdiv(b.n(), 256), 256, 0, stream()>>>(E_ARGS, INT_ARGS);
        else yee_E_int_kernel<false><<<cdiv(b.n(), 256), 256, 0, stream()>>>(E_ARGS, INT_ARGS);
        MP_LAUNCH_CHECK("yee_E_int_kernel");
    }
    // the whole shell in one launch: split like the H shell (x strips, the box's x range of the
    // rest rows, their x edges) it measured 682 against 666 us in the 512 x 512 x 256 pump step
    // and 611 against 590 us in the 384^3 stencil
    const unsigned grid = cdiv(sg.n, 64 * TY);
    if (mi)
        yee_E_vec_kernel<P, TY, true, true><<<grid, block, 0, stream()>>>(
            E_ARGS, make_cpml_args(c, false), nx, ny, nz, 1, (float)dt, 1, sg);
    else
        yee_E_vec_kernel<P, TY, false, true><<<grid, block, 0, stream()>>>(
            E_ARGS, make_cpml_args(c, false), nx, ny, nz, 1, (float)dt, 1, sg);
    MP_LAUNCH_CHECK("yee_E_vec_kernel");
#undef INT_ARGS
#undef E_ARGS
    return MP_OK;
}

template <typename P>
static int launch_H_vec(const mp_yee *g, const mp_cpml *c, double dt, double mu0) {
    return launch_H_vec_t<P, TYV>(g, c, dt, mu0);
}

template <typename P>
static int launch_E_vec(const mp_yee *g, const mp_cpml *c, double dt) {
    return launch_E_vec_t<P, TYV>(g, c, dt);
}

// The vector form needs x rows of a multiple of 4 cells, the volumes and coefficient lines it
// reads as float4 on 16-byte boundaries, and the y/z-slab psi arrays on boundaries of four
// of their elements; anything else takes the scalar kernels.
static bool al(const void *p, size_t a) { return (uintptr_t)p % a == 0; }
static bool vec_ok(const mp_yee *g, const mp_cpml *c, bool h) {
    if (g->kernel_form == MP_YEE_SCALAR || g->nx % VX != 0) return false;
    const void *v16[] = {g->ex, g->ey, g->ez, g->hx, g->hy, g->hz, g->dx, g->dy, g->dz,
                         g->inv_eps_x, g->inv_eps_y, g->inv_eps_z, g->inv_d_cell_x, g->inv_d_dual_x,
                         c->inv_kappa[0]};
    for (const void *p : v16)
        if (!al(p, 16)) return false;
    if (!al(g->eps_index, 4)) return false;
    const size_t q = c->psi_dtype == MP_F32 ? 16 : 8;
    const int yz[] = {0, 1, 2, 3, 4, 5, 10, 11};
    for (int s : yz)
        if (!al(h ? c->psi_h[s] : c->psi_d[s], q)) return false;
    return true;
}

template <typename S, typename C, typename P>
static int launch_H(const mp_yee *g, const mp_cpml *c, double dt, double mu0) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    const int64_t kc = pick_kchunk(nx, ny, nz);
    dim3 grid(cdiv(nx, TX), cdiv(ny, TY), cdiv(nz, kc));
    dim3 block(TX, TY, 1);
    const C dt_mu0 = (C)dt / (C)mu0;
    KernelTimer t("yee_H");
    yee_H_kernel<S, C, P><<<grid, block, 0, stream()>>>(
        (const S *)g->ex, (const S *)g->ey, (const S *)g->ez, (S *)g->hx, (S *)g->hy, (S *)g->hz,
        (const C *)g->inv_d_cell_x, (const C *)g->inv_d_cell_y, (const C *)g->inv_d_cell_z,
        make_cpml_args(c, true), nx, ny, nz, kc, dt_mu0);
    MP_LAUNCH_CHECK("yee_H_kernel");
    return MP_OK;
}

template <typename S, typename C, typename P>
static int launch_E(const mp_yee *g, const mp_cpml *c, double dt) {
    const int64_t nx = g->nx, ny = g->ny, nz = g->nz;
    const int64_t kc = pick_kchunk(nx, ny, nz);
    dim3 grid(cdiv(nx, TX), cdiv(ny, TY), cdiv(nz, kc));
    dim3 block(TX, TY, 1);
    KernelTimer t("yee_E");
    yee_E_kernel<S, C, P><<<grid, block, 0, stream()>>>(
        (S *)g->ex, (S *)g->ey, (S *)g->ez, (const S *)g->hx, (const S *)g->hy, (const S *)g->hz,
        (S *)g->dx, (S *)g->dy, (S *)g->dz, (const S *)g->inv_eps_x, (const S *)g->inv_eps_y,
        (const S *)g->inv_eps_z, (const C *)g->inv_d_dual_x, (const C *)g->inv_d_dual_y,
        (const C *)g->inv_d_dual_z, make_cpml_args(c, false), nx, ny, nz, kc, (C)dt);
    MP_LAUNCH_CHECK("yee_E_kernel");
    return MP_OK;
}

// ---- ping-pong steps for the step driver -------------------------------------------------
bool yee_fused_ok(const mp_yee *g, const mp_cpml *c) {
    if (!g || !c || check_yee(g, c, "yee_fused_ok") != MP_OK) return false;
    if (g->field_dtype != MP_F32 || g->compute_dtype != MP_F32 || g->kernel_form != MP_YEE_AUTO) return false;
    if (!(g->nx > 1 && g->ny > 1 && g->nz > 1)) return false;
    if (c->psi_dtype != MP_F32 && c->psi_dtype != MP_BF16 && c->psi_dtype != MP_F16) return false;
    if (!vec_ok(g, c, true) || !vec_ok(g, c, false)) return false;
    const Box b = interior_box(g, c);
    if (!split_ok(g, b) || !shell_geom(g, b).n) return false;
    // the recomputed row of a workgroup (ngi lanes) must leave room for as many owned lanes;
    // the fused kernel addresses the volumes by 32-bit byte offsets
    return b.ngi >= 1 && 2 * b.ngi <= FL && 4 * g->nx * g->ny * g->nz < ((int64_t)1 << 32);
}

int yee_fused_H_shell(const mp_yee *in, const mp_yee *out, mp_cpml *c, double dt, double mu0) {
    switch (c->psi_dtype) {
    case MP_F32: return launch_H_vec_t<float, TYV>(in, c, dt, mu0, out);
    case MP_BF16: return launch_H_vec_t<bf16, TYV>(in, c, dt, mu0, out);
    case MP_F16: return launch_H_vec_t<f16s, TYV>(in, c, dt, mu0, out);
    default: return fail_arg("fused Yee step: unsupported psi dtype %d", c->psi_dtype);
    }
}

int yee_fused_E_shell(const mp_yee *out, mp_cpml *c, double dt) {
    switch (c->psi_dtype) {
    case MP_F32: return launch_E_vec_t<float, TYV>(out, c, dt, true);
    case MP_BF16: return launch_E_vec_t<bf16, TYV>(out, c, dt, true);
    case MP_F16: return launch_E_vec_t<f16s, TYV>(out, c, dt, true);
    default: return fail_arg("fused Yee step: unsupported psi dtype %d", c->psi_dtype);
    }
}

int yee_fused_interior(const mp_yee *in, const mp_yee *out, const mp_cpml *c, double dt, double mu0) {
    const Box b = interior_box(in, c);
    const int64_t nplane = b.ngi * b.nj, T = FL - b.ngi;
    const int64_t tiles = cdiv(nplane, T);
    // about 1024 workgroups (two to four per CU), planes per chunk at least 8
    const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(cdiv(1024, tiles), b.nk / 8));
    const int64_t kchunk = cdiv(b.nk, chunks);
    const int64_t nwg = tiles * cdiv(b.nk, kchunk);
    if (nwg >= ((int64_t)1 << 31)) return fail_arg("fused Yee step: grid too large");
    const bool mi = in->eps_index && in->eps_table && in->n_eps > 0 && in->n_eps <= 256;
    const float dt_mu0 = (float)dt / (float)mu0;
    KernelTimer t("yee_HE");
#define HE_ARGS                                                                                                \
    (const float *)in->ex, (const float *)in->ey, (const float *)in->ez, (const float *)in->hx,                \
        (const float *)in->hy, (const float *)in->hz, (float *)out->ex, (float *)out->ey, (float *)out->ez,    \
        (float *)out->hx, (float *)out->hy, (float *)out->hz, (float *)in->dx, (float *)in->dy,                 \
        (float *)in->dz, (const float *)in->inv_eps_x, (const float *)in->inv_eps_y,                            \
        (const float *)in->inv_eps_z, in->eps_index, (const float4 *)in->eps_table, in->n_eps,                 \
        (const float *)in->inv_d_cell_x, (const float *)in->inv_d_cell_y, (const float *)in->inv_d_cell_z,     \
        (const float *)in->inv_d_dual_x, (const float *)in->inv_d_dual_y, (const float *)in->inv_d_dual_z,     \
        (const float *)c->inv_kappa[0], (const float *)c->inv_kappa[1], (const float *)c->inv_kappa[2],        \
        (uint32_t)in->nx, (uint32_t)in->ny, (uint32_t)b.gi0, (uint32_t)b.ngi, (uint32_t)b.j0, (uint32_t)nplane, \
        (uint32_t)b.k0, (uint32_t)(b.k0 + b.nk), (uint32_t)tiles, (uint32_t)kchunk, dt_mu0, (float)dt
    if (mi) yee_HE_int_kernel<true><<<(unsigned)nwg, FL, 0, stream()>>>(HE_ARGS);
    else yee_HE_int_kernel<false><<<(unsigned)nwg, FL, 0, stream()>>>(HE_ARGS);
#undef HE_ARGS
    MP_LAUNCH_CHECK("yee_HE_int_kernel");
    return MP_OK;
}

int yee_sync_boundary(const mp_yee *from, const mp_yee *to, bool e_only) {
    const int64_t nx = from->nx, ny = from->ny, nz = from->nz;
    const PlanePtrs a{{(float *)from->ex, (float *)from->ey, (float *)from->ez, (float *)from->hx,
                       (float *)from->hy, (float *)from->hz}};
    const PlanePtrs bb{{(float *)to->ex, (float *)to->ey, (float *)to->ez, (float *)to->hx, (float *)to->hy,
                        (float *)to->hz}};
    const int64_t n = std::max(ny * nz, std::max(nx * nz, nx * ny));
    KernelTimer t("pingpong_sync");
    sync_boundary_kernel<<<dim3(cdiv(n, 256), e_only ? 3 : 6), 256, 0, stream()>>>(a, bb, nx, ny, nz);
    MP_LAUNCH_CHECK("sync_boundary_kernel");
    return MP_OK;
}

}  // namespace mp

using namespace mp;

MP_API int mp_update_H_3d(const mp_yee *g, mp_cpml *c, double dt, double mu0) {
    int rc = check_yee(g, c, "mp_update_H_3d");
    if (rc) return rc;
    if (!(g->nx > 1 && g->ny > 1 && g->nz > 1)) return MP_OK;  // Kernels.jl:528
    const int32_t fd = g->field_dtype, cd = g->compute_dtype, pd = c->psi_dtype;
    if (fd == MP_F32 && cd == MP_F32 && vec_ok(g, c, true)) {
        if (pd == MP_F32) return launch_H_vec<float>(g, c, dt, mu0);
        if (pd == MP_BF16) return launch_H_vec<bf16>(g, c, dt, mu0);
        if (pd == MP_F16) return launch_H_vec<f16s>(g, c, dt, mu0);
    } else if (fd == MP_F32 && cd == MP_F32) {
        if (pd == MP_F32) return launch_H<float, float, float>(g, c, dt, mu0);
        if (pd == MP_BF16) return launch_H<float, float, bf16>(g, c, dt, mu0);
        if (pd == MP_F16) return launch_H<float, float, f16s>(g, c, dt, mu0);
    } else if (fd == MP_F64 && cd == MP_F64) {
        if (pd == MP_F64) return launch_H<double, double, double>(g, c, dt, mu0);
        if (pd == MP_F32) return launch_H<double, double, float>(g, c, dt, mu0);
    } else if (fd == MP_F32 && cd == MP_F64) {
        if (pd == MP_F32) return launch_H<float, double, float>(g, c, dt, mu0);
    }
    return fail_arg("mp_update_H_3d: unsupported dtypes field=%d compute=%d psi=%d", fd, cd, pd);
}

MP_API int mp_update_E_3d(const mp_yee *g, mp_cpml *c, double dt) {
    int rc = check_yee(g, c, "mp_update_E_3d");
    if (rc) return rc;
    if (!(g->nx > 1 && g->ny > 1 && g->nz > 1)) return MP_OK;  // Kernels.jl:550
    const int32_t fd = g->field_dtype, cd = g->compute_dtype, pd = c->psi_dtype;
    if (fd == MP_F32 && cd == MP_F32 && vec_ok(g, c, false)) {
        if (pd == MP_F32) return launch_E_vec<float>(g, c, dt);
        if (pd == MP_BF16) return launch_E_vec<bf16>(g, c, dt);
        if (pd == MP_F16) return launch_E_vec<f16s>(g, c, dt);
    } else if (fd == MP_F32 && cd == MP_F32) {
        if (pd == MP_F32) return launch_E<float, float, float>(g, c, dt);
        if (pd == MP_BF16) return launch_E<float, float, bf16>(g, c, dt);
        if (pd == MP_F16) return launch_E<float, float, f16s>(g, c, dt);
    } else if (fd == MP_F64 && cd == MP_F64) {
        if (pd == MP_F64) return launch_E<double, double, double>(g, c, dt);
        if (pd == MP_F32) return launch_E<double, double, float>(g, c, dt);
    } else if (fd == MP_F32 && cd == MP_F64) {
        if (pd == MP_F32) return launch_E<float, double, float>(g, c, dt);
    }
    return fail_arg("mp_update_E_3d: unsupported dtypes field=%d compute=%d psi=%d", fd, cd, pd);
}
