// Library-level entry points: error reporting, device init, scene sizes.
#include "gf_common.h"
#include <stdarg.h>
#include <stdio.h>

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

}  // namespace gf

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
    out->partials = 8 * 2048;  // per-workgroup partial sums of the PCG reductions
    out->scalars = 64;         // two 16-double solver log entries, plus spare
    return 0;
}
