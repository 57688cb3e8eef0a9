// Movie frames of the HDF5 writers staged on the device: |field| of one plane of a field
// volume as float32, the host side's Float32.(abs.(view)) of _stage_abs_frame!
// (ext/MagnetoPhotonicHDF5Ext.jl:619-624), which record_production_frame! (:626-636) and
// record_probe_frame! (:685-697) write frame by frame.
//
// Kept in a translation unit of its own: the step's time is sensitive to the layout of the
// monitor kernels' code object (adding this kernel to monitors.hip slowed the 512 x 512 x
// 256 pump step from 1.487 to 1.52-1.53 ms, reproducibly on one box; the film update and
// the H update beside it both ran slower).
#include "common.h"

namespace mp {

// |field| of one plane as float, in the plane order of plane_gather_kernel.
template <typename T>
__global__ __launch_bounds__(256) void plane_abs_kernel(const T *f, int axis, int64_t idx, int64_t nx, int64_t ny,
                                                        int64_t n1, int64_t n2, float *out) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n1 * n2) return;
    const int64_t a = c % n1, b = c / n1;
    int64_t o;
    if (axis == 0) o = idx + nx * (a + ny * b);
    else if (axis == 1) o = a + nx * (idx + ny * b);
    else o = a + nx * (b + ny * idx);
    out[c] = (float)__builtin_fabs(f[o]);
}

}  // namespace mp

using namespace mp;

MP_API int mp_plane_abs(const void *field, int32_t dtype, int64_t nx, int64_t ny, int64_t nz, int32_t axis,
                        int64_t index, float *out) {
    if (!field || !out) return fail_arg("mp_plane_abs: null argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail_arg("mp_plane_abs: bad dims");
    if (axis < 0 || axis > 2) return fail_arg("3-D field plane axis must be :x, :y, or :z");
    const int64_t nax = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if (index < 0 || index >= nax) return fail_arg("mp_plane_abs: plane index out of range");
    const int64_t n1 = axis == 0 ? ny : nx, n2 = axis == 2 ? ny : nz;
    KernelTimer t("plane_abs");
    if (dtype == MP_F32)
        plane_abs_kernel<float><<<cdiv(n1 * n2, 256), 256, 0, stream()>>>((const float *)field, axis, index, nx, ny,
                                                                          n1, n2, out);
    else if (dtype == MP_F64)
        plane_abs_kernel<double><<<cdiv(n1 * n2, 256), 256, 0, stream()>>>((const double *)field, axis, index, nx, ny,
                                                                           n1, n2, out);
    else
        return fail_arg("mp_plane_abs: unsupported dtype %d", dtype);
    MP_LAUNCH_CHECK("plane_abs_kernel");
    return MP_OK;
}
