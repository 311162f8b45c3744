// Library-level entry points: error reporting, device init, scene sizes.
#include "gf_common.h"
#include <atomic>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

namespace gf {

static thread_local char g_err[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error((int)e, "%s: %s", what, hipGetErrorString(e));
        return (int)e;
    }
    return 0;
}

// Launch-shape tunables (0 = the built-in choice).  Read at every call.
static std::atomic<int> g_tunables[GF_TUNE_COUNT];

int tunable(int id) { return g_tunables[id].load(std::memory_order_relaxed); }

}  // namespace gf

static int tune_id(const char* key) {
    if (!key) return -1;
    if (!strcmp(key, "pcg_workgroups")) return gf::GF_TUNE_PCG_WORKGROUPS;
    if (!strcmp(key, "pcg_variant")) return gf::GF_TUNE_PCG_VARIANT;
    if (!strcmp(key, "pcg_x_every")) return gf::GF_TUNE_PCG_X_EVERY;
    return -1;
}

extern "C" int gf_tune_set(const char* key, int value) {
    const int id = tune_id(key);
    if (id < 0) return gf::set_error(GF_EINVAL, "gf_tune_set: unknown key '%s'", key ? key : "(null)");
    if (value < 0) return gf::set_error(GF_EINVAL, "gf_tune_set: %s = %d", key, value);
    gf::g_tunables[id].store(value, std::memory_order_relaxed);
    return 0;
}

extern "C" int gf_tune_get(const char* key, int* value) {
    const int id = tune_id(key);
    if (id < 0 || !value)
        return gf::set_error(GF_EINVAL, "gf_tune_get: unknown key '%s'", key ? key : "(null)");
    *value = gf::tunable(id);
    return 0;
}

extern "C" int gf_version(void) { return 10000; }

extern "C" const char* gf_last_error(void) { return gf::g_err; }

extern "C" int gf_init(int device, int* cus) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess)
        return gf::set_error((int)e, "gf_init: hipSetDevice(%d): %s", device,
                             hipGetErrorString(e));
    if (cus) {
        int v = 0;
        e = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess)
            return gf::set_error((int)e, "gf_init: attribute query: %s", hipGetErrorString(e));
        *cus = v;
    }
    return 0;
}

extern "C" int gf_scene_sizes(int nx, int ny, int nz, gf_scene_size_info* out) {
    if (!out) return gf::set_error(GF_EINVAL, "gf_scene_sizes: null output");
    if (nx < 1 || ny < 1 || nz < 1)
        return gf::set_error(GF_EINVAL, "gf_scene_sizes: grid %dx%dx%d", nx, ny, nz);
    int64_t cells = (int64_t)nx * ny * nz;
    if (cells + 1 > (int64_t)INT32_MAX)
        return gf::set_error(GF_ERANGE, "gf_scene_sizes: %lld cells exceed int32 indexing",
                             (long long)cells);
    int64_t ntiles = (int64_t)((nx + 7) / 8) * ((ny + 7) / 8) * ((nz + 7) / 8);
    out->cells = cells;
    out->ntiles = ntiles;
    out->compact = ntiles * GF_TILE_CELLS;
    out->scan_parts = (cells + 1 + 4095) / 4096 + 1;
    out->partials = 2 * 8 * 2048;  // per-workgroup PCG partial sums, two launch-parity sets
    out->scalars = 64;         // three 16-double solver log entries (two live, one result), spare
    out->pcg_part = 8 * 2048 + 2;  // per-workgroup scene/tile ranges + per-scene offsets
    return 0;
}
