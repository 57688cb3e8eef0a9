// libmphip runtime: device selection, memory, copies, stream, events, per-kernel timing
// and error strings.  These are the backend hooks of src/core/Backend.jl:19-126
// (zeros_backend / adapt_backend / to_host / synchronize / device_memory_info /
// free_device!) that a HIP package extension must provide.
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "monitors.h"

namespace mp {

static thread_local char g_err[512] = {0};
static hipStream_t g_stream = nullptr;
static bool g_profile = false;
static std::vector<std::string> g_profile_only;   // empty: every kernel name is timed
static int g_profile_every = 1;                    // time every n-th launch of each name
static std::map<std::string, int64_t> g_profile_seen;

struct PendingTiming {
    const char *name;
    hipEvent_t start, stop;
};
static std::vector<PendingTiming> g_pending;
static std::map<std::string, std::pair<int64_t, double>> g_stats;
static std::vector<hipEvent_t> g_event_pool;

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail_arg(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return MP_ERR_ARG;
}

int check_hip(hipError_t e, const char *what) {
    if (e == hipSuccess) return MP_OK;
    set_error("%s: %s", what, hipGetErrorString(e));
    return MP_ERR_HIP;
}

hipStream_t stream() { return g_stream; }

hipStream_t swap_stream(hipStream_t s) {
    hipStream_t old = g_stream;
    g_stream = s;
    return old;
}

static hipEvent_t pool_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static void drain_pending() {
    for (auto &p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto &s = g_stats[p.name];
            s.first += 1;
            s.second += ms;
        }
        g_event_pool.push_back(p.start);
        g_event_pool.push_back(p.stop);
    }
    g_pending.clear();
}

static bool profiled(const char *name) {
    if (g_profile_only.empty()) return true;
    for (const std::string &n : g_profile_only)
        if (n == name) return true;
    return false;
}

KernelTimer::KernelTimer(const char *name) : name_(name) {
    if (!g_profile || !profiled(name)) return;
    if (g_profile_every > 1 && g_profile_seen[name]++ % g_profile_every != 0) return;
    hipEvent_t e = pool_event();
    if (e && hipEventRecord(e, g_stream) == hipSuccess) start_ = (void *)e;
}

KernelTimer::~KernelTimer() {
    if (!start_) return;
    hipEvent_t stop = pool_event();
    if (stop && hipEventRecord(stop, g_stream) == hipSuccess) {
        g_pending.push_back({name_, (hipEvent_t)start_, stop});
        if (g_pending.size() > 4096) drain_pending();
    }
}

}  // namespace mp

using namespace mp;

MP_API int mp_init(int device) {
    int n = 0;
    MP_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail_arg("mp_init: device %d out of range (%d devices)", device, n);
    MP_HIP(hipSetDevice(device));
    MP_HIP(hipFree(nullptr));  // create the context now
    double *scratch_dev = nullptr;
    return noalloc_scratch(&scratch_dev);  // this device's fixed scratch, resolved once
}

MP_API const char *mp_last_error(void) { return g_err; }

MP_API int mp_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail_arg("mp_malloc: null out pointer");
    *ptr = nullptr;
    if (bytes == 0) return MP_OK;
    MP_HIP(hipMalloc(ptr, bytes));
    return MP_OK;
}

MP_API int mp_free(void *ptr) {
    if (!ptr) return MP_OK;
    MP_HIP(hipFree(ptr));
    return MP_OK;
}

MP_API int mp_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MP_OK;
    MP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
    MP_HIP(hipStreamSynchronize(g_stream));
    return MP_OK;
}

MP_API int mp_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MP_OK;
    MP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
    MP_HIP(hipStreamSynchronize(g_stream));
    return MP_OK;
}

MP_API int mp_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return MP_OK;
    MP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
    return MP_OK;
}

MP_API int mp_memset(void *ptr, int value, size_t bytes) {
    if (bytes == 0) return MP_OK;
    MP_HIP(hipMemsetAsync(ptr, value, bytes, g_stream));
    return MP_OK;
}

template <typename T>
__global__ void fill_kernel(T *p, int64_t n, T v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

MP_API int mp_fill(void *ptr, int32_t dtype, int64_t n, double value) {
    if (n <= 0) return MP_OK;
    if (!ptr) return fail_arg("mp_fill: null pointer");
    const unsigned blocks = (unsigned)std::min<int64_t>(cdiv(n, 256), 65535);
    KernelTimer t("fill");
    if (dtype == MP_F32) {
        fill_kernel<float><<<blocks, 256, 0, g_stream>>>((float *)ptr, n, (float)value);
    } else if (dtype == MP_F64) {
        fill_kernel<double><<<blocks, 256, 0, g_stream>>>((double *)ptr, n, value);
    } else {
        return fail_arg("mp_fill: unsupported dtype %d", dtype);
    }
    MP_LAUNCH_CHECK("fill_kernel");
    return MP_OK;
}

MP_API int mp_synchronize(void) {
    MP_HIP(hipStreamSynchronize(g_stream));
    return MP_OK;
}

MP_API int mp_mem_info(size_t *free_bytes, size_t *total_bytes) {
    MP_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return MP_OK;
}

MP_API int mp_set_stream(void *s) {
    g_stream = (hipStream_t)s;
    return MP_OK;
}

MP_API void *mp_get_stream(void) { return (void *)g_stream; }

MP_API int mp_event_create(void **event) {
    hipEvent_t e = nullptr;
    MP_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return MP_OK;
}

MP_API int mp_event_destroy(void *event) {
    if (event) MP_HIP(hipEventDestroy((hipEvent_t)event));
    return MP_OK;
}

MP_API int mp_event_record(void *event) {
    MP_HIP(hipEventRecord((hipEvent_t)event, g_stream));
    return MP_OK;
}

MP_API int mp_event_elapsed_ms(void *start, void *stop, float *ms) {
    MP_HIP(hipEventSynchronize((hipEvent_t)stop));
    MP_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return MP_OK;
}

MP_API int mp_profile_enable(int on) {
    drain_pending();
    if (on) g_stats.clear();
    g_profile = on != 0;
    return MP_OK;
}

MP_API int mp_profile_select(const char *names) {
    drain_pending();
    g_profile_only.clear();
    if (!names) return MP_OK;
    std::string cur;
    for (const char *c = names;; ++c) {
        if (*c == ',' || *c == '\0') {
            if (!cur.empty()) g_profile_only.push_back(cur);
            cur.clear();
            if (*c == '\0') break;
        } else {
            cur.push_back(*c);
        }
    }
    return MP_OK;
}

MP_API int mp_profile_sample(int32_t every) {
    if (every < 1) return fail_arg("mp_profile_sample: every must be >= 1 (got %d)", every);
    drain_pending();
    g_profile_every = every;
    g_profile_seen.clear();
    return MP_OK;
}

MP_API int mp_kernel_stats(const char *name, int64_t *launches, double *total_ms) {
    drain_pending();
    auto it = g_stats.find(name ? name : "");
    if (it == g_stats.end()) {
        *launches = 0;
        *total_ms = 0.0;
        return MP_OK;
    }
    *launches = it->second.first;
    *total_ms = it->second.second;
    return MP_OK;
}
