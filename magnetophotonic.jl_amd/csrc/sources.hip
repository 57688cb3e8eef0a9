// Source injection kernels.
//
// Reference: inject_soft_3d_kernel! (src/fdtd/Kernels.jl:154-174), inject_mode_x_3d_kernel!
// (:176-192) and the host loop inject!(::FieldState, ::Grid3D, ::PlaneSource)
// (src/fdtd/Source.jl:188-236), which the reference runs on the CPU only.
#include "common.h"
#include "sources.h"

namespace mp {

// E (or H) is read from Ein (Hin) and written to E (H): the same array for the entry points,
// the other buffer set for the step driver's ping-pong step.  D is updated in place.
template <typename S, typename C>
__global__ void inject_soft_kernel(S *E, const S *Ein, S *D, const S *ie, S *H, const S *Hin, int64_t o, C val) {
    if (E) {
        E[o] = (S)((C)Ein[o] + val);
        D[o] = (S)((C)D[o] + val / (C)ie[o]);
    } else {
        H[o] = (S)((C)Hin[o] + val);
    }
}

template <typename S, typename C>
__global__ void inject_mode_x_kernel(S *E, S *D, const S *ie, S *H, const C *profile, int64_t i,
                                     int64_t nx, int64_t ny, int64_t nz, C value, C hsign_hcorr) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t k = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
    if (j >= ny || k >= nz) return;
    const int64_t o = i + nx * (j + ny * k);
    const C amp = value * profile[j + ny * k];
    E[o] = (S)((C)E[o] + amp);
    D[o] = (S)((C)D[o] + amp / (C)ie[o]);
    // Ey source: Hz += hcorr*amp; Ez source: Hy -= hcorr*amp (Kernels.jl:186,190)
    H[o] = (S)((C)H[o] + hsign_hcorr * amp);
}

// Plane injection: E[c] = S(double(E[c]) + dv), D[c] = S(double(D[c]) + dv / double(ie[c]))
// over every cell of plane `idx` of `axis` (Source.jl:216-236 promotes to Float64).
// blockIdx.z selects one of the two interpolation planes (distinct planes: no cell is
// touched twice within a launch).
template <typename S>
__global__ void inject_plane_kernel(S *F, const S *Fin, S *D, const S *ie, int axis, int64_t idx0, double dv0,
                                    int64_t idx1, double dv1, int64_t nx, int64_t ny, int64_t nz) {
    const int64_t idx = blockIdx.z ? idx1 : idx0;
    const double dv = blockIdx.z ? dv1 : dv0;
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
    int64_t o;
    if (axis == 0) {        // plane (j, k)
        if (a >= ny || b >= nz) return;
        o = idx + nx * (a + ny * b);
    } else if (axis == 1) {  // plane (i, k)
        if (a >= nx || b >= nz) return;
        o = a + nx * (idx + ny * b);
    } else {                 // plane (i, j)
        if (a >= nx || b >= ny) return;
        o = a + nx * (b + ny * idx);
    }
    F[o] = (S)((double)Fin[o] + dv);
    if (D) D[o] = (S)((double)D[o] + dv / (double)ie[o]);
}

struct CompPtrs {
    void *E, *D, *H;
    const void *ie;
};

static CompPtrs comp_ptrs(const mp_yee *g, int code) {
    switch (code) {
    case 1: return {g->ex, g->dx, nullptr, g->inv_eps_x};
    case 2: return {g->ey, g->dy, nullptr, g->inv_eps_y};
    case 3: return {g->ez, g->dz, nullptr, g->inv_eps_z};
    case 4: return {nullptr, nullptr, g->hx, nullptr};
    case 5: return {nullptr, nullptr, g->hy, nullptr};
    case 6: return {nullptr, nullptr, g->hz, nullptr};
    default: return {nullptr, nullptr, nullptr, nullptr};
    }
}

}  // namespace mp

using namespace mp;

namespace mp {

int inject_soft_pp(const mp_yee *g, const mp_yee *out, int32_t component, int64_t i, int64_t j, int64_t k,
                   double value) {
    if (!g || !out) return fail_arg("mp_inject_soft_3d: null grid");
    if (component < 1 || component > 6) return fail_arg("mp_inject_soft_3d: bad component %d", component);
    if (i < 0 || j < 0 || k < 0 || i >= g->nx || j >= g->ny || k >= g->nz)
        return fail_arg("mp_inject_soft_3d: index (%lld,%lld,%lld) out of range", (long long)i,
                        (long long)j, (long long)k);
    const int64_t o = i + g->nx * (j + g->ny * k);
    const CompPtrs p = comp_ptrs(g, component), q = comp_ptrs(out, component);
    KernelTimer t("inject_soft");
    if (g->field_dtype == MP_F32 && g->compute_dtype == MP_F32) {
        inject_soft_kernel<float, float><<<1, 1, 0, stream()>>>((float *)q.E, (const float *)p.E, (float *)p.D,
            (const float *)p.ie, (float *)q.H, (const float *)p.H, o, (float)value);
    } else if (g->field_dtype == MP_F64 && g->compute_dtype == MP_F64) {
        inject_soft_kernel<double, double><<<1, 1, 0, stream()>>>((double *)q.E, (const double *)p.E,
            (double *)p.D, (const double *)p.ie, (double *)q.H, (const double *)p.H, o, value);
    } else if (g->field_dtype == MP_F32 && g->compute_dtype == MP_F64) {
        inject_soft_kernel<float, double><<<1, 1, 0, stream()>>>((float *)q.E, (const float *)p.E, (float *)p.D,
            (const float *)p.ie, (float *)q.H, (const float *)p.H, o, value);
    } else {
        return fail_arg("mp_inject_soft_3d: unsupported dtypes");
    }
    MP_LAUNCH_CHECK("inject_soft_kernel");
    return MP_OK;
}

}  // namespace mp

MP_API int mp_inject_soft_3d(const mp_yee *g, int32_t component, int64_t i, int64_t j, int64_t k,
                             double value) {
    return inject_soft_pp(g, g, component, i, j, k, value);
}

MP_API int mp_inject_mode_x_3d(const mp_yee *g, int32_t component, int64_t i, const void *profile,
                               double value, double hcorr) {
    if (!g || !profile) return fail_arg("mp_inject_mode_x_3d: null argument");
    if (component != MP_EY && component != MP_EZ)
        return fail_arg("mp_inject_mode_x_3d: ModeSource electric component must be Ey or Ez");
    if (i < 0 || i >= g->nx) return fail_arg("mp_inject_mode_x_3d: plane index out of range");
    void *E = component == MP_EY ? g->ey : g->ez;
    void *D = component == MP_EY ? g->dy : g->dz;
    const void *ie = component == MP_EY ? g->inv_eps_y : g->inv_eps_z;
    void *H = component == MP_EY ? g->hz : g->hy;
    dim3 block(64, 4);
    dim3 grid(cdiv(g->ny, 64), cdiv(g->nz, 4));
    KernelTimer t("inject_mode_x");
    if (g->field_dtype == MP_F32 && g->compute_dtype == MP_F32) {
        const float hc = component == MP_EY ? (float)hcorr : -(float)hcorr;
        inject_mode_x_kernel<float, float><<<grid, block, 0, stream()>>>((float *)E, (float *)D,
            (const float *)ie, (float *)H, (const float *)profile, i, g->nx, g->ny, g->nz,
            (float)value, hc);
    } else if (g->field_dtype == MP_F64 && g->compute_dtype == MP_F64) {
        const double hc = component == MP_EY ? hcorr : -hcorr;
        inject_mode_x_kernel<double, double><<<grid, block, 0, stream()>>>((double *)E, (double *)D,
            (const double *)ie, (double *)H, (const double *)profile, i, g->nx, g->ny, g->nz, value, hc);
    } else if (g->field_dtype == MP_F32 && g->compute_dtype == MP_F64) {
        const double hc = component == MP_EY ? hcorr : -hcorr;
        inject_mode_x_kernel<float, double><<<grid, block, 0, stream()>>>((float *)E, (float *)D,
            (const float *)ie, (float *)H, (const double *)profile, i, g->nx, g->ny, g->nz, value, hc);
    } else {
        return fail_arg("mp_inject_mode_x_3d: unsupported dtypes");
    }
    MP_LAUNCH_CHECK("inject_mode_x_kernel");
    return MP_OK;
}

namespace mp {

int inject_plane_pp(const mp_yee *g, const mp_yee *out, int32_t component, int32_t axis, int64_t i0, double w0,
                    int64_t i1, double w1, double value) {
    if (!g || !out) return fail_arg("mp_inject_plane_3d: null grid");
    if (component < 1 || component > 6) return fail_arg("mp_inject_plane_3d: bad component %d", component);
    if (axis < 0 || axis > 2) return fail_arg("mp_inject_plane_3d: 3-D PlaneSource axis must be x, y or z");
    const int64_t n_ax = axis == 0 ? g->nx : (axis == 1 ? g->ny : g->nz);
    if (i0 < 0 || i0 >= n_ax || i1 >= n_ax) return fail_arg("mp_inject_plane_3d: plane index out of range");
    const CompPtrs p = comp_ptrs(g, component), q = comp_ptrs(out, component);
    void *Fin = p.E ? p.E : p.H;
    void *F = q.E ? q.E : q.H;
    const int64_t n1 = axis == 0 ? g->ny : g->nx;
    const int64_t n2 = axis == 2 ? g->ny : g->nz;
    if (g->field_dtype != MP_F32 && g->field_dtype != MP_F64)
        return fail_arg("mp_inject_plane_3d: unsupported field dtype");
    dim3 block(64, 4);
    const double dv0 = w0 * value, dv1 = w1 * value;
    // both planes in one launch when they differ; the same plane twice keeps the order (the
    // second pass adds to what the first wrote)
    const bool both = i1 >= 0 && i1 != i0;
    const int passes = (i1 >= 0 && !both) ? 2 : 1;
    KernelTimer t("inject_plane");
    for (int s = 0; s < passes; ++s) {
        dim3 grid(cdiv(n1, 64), cdiv(n2, 4), both ? 2 : 1);
        const double d0 = s == 0 ? dv0 : dv1;
        const void *src = s == 0 ? Fin : F;
        if (g->field_dtype == MP_F32)
            inject_plane_kernel<float><<<grid, block, 0, stream()>>>((float *)F, (const float *)src, (float *)p.D,
                                                                     (const float *)p.ie, axis, i0, d0, i1, dv1,
                                                                     g->nx, g->ny, g->nz);
        else
            inject_plane_kernel<double><<<grid, block, 0, stream()>>>((double *)F, (const double *)src,
                                                                      (double *)p.D, (const double *)p.ie, axis,
                                                                      i0, d0, i1, dv1, g->nx, g->ny, g->nz);
        MP_LAUNCH_CHECK("inject_plane_kernel");
    }
    return MP_OK;
}

}  // namespace mp

MP_API int mp_inject_plane_3d(const mp_yee *g, int32_t component, int32_t axis, int64_t i0,
                              double w0, int64_t i1, double w1, double value) {
    return inject_plane_pp(g, g, component, axis, i0, w0, i1, w1, value);
}
