// Solid cells from a closed triangle mesh: the inside test of the mesh
// obstacles (D4.3.GPU, schemes/mesh_marker.py:39-80, 197-236) that
// prepare_frame applies to animated mesh obstacles (solver3d.py:2598-2631).
//
// A cell is inside when the ray from its FP32 centre ((i + 0.5) * dx rounded
// once) along +x crosses the mesh an odd number of times; each crossing is the
// Moller-Trumbore test with direction (1, 0, 0), in the reference's operation
// order with every operation rounded on its own.  With that direction the
// barycentric coordinates depend only on the ray's (y, z), so a triangle can
// only meet the rays of the grid columns (j, k) whose centre (y, z) lies in
// its y-z bounding box.  The triangles are therefore binned into columns
// first (count, scan, fill), with the box widened by a margin far above the
// rounding error of the barycentric test, and each cell tests its column's
// triangles only.  A crossing count's parity does not depend on the order in
// which the triangles are visited, so the binning changes no result.
#include "gf_common.h"

namespace {

struct MeshArgs {
    const float* tris;  // n_tris * 9: v0, v1, v2
    int n_tris, nx, ny, nz;
    double dx;
    double extent;      // largest domain edge
};

struct MeshWs {
    int32_t* counts;  // [cols] triangles per column
    int32_t* start;   // [cols + 1] exclusive prefix
    int32_t* cursor;  // [cols] fill cursors
    int32_t* part;    // [scan parts]
    unsigned long long* total;  // 64-bit sum of the column ranges
};

int64_t scan_parts(int64_t cols) { return (cols + 4095) / 4096 + 1; }

int64_t mesh_ws_ints(int64_t cols) { return 3 * cols + 1 + scan_parts(cols) + 4; }

MeshWs mesh_ws(int32_t* ws, int64_t cols) {
    MeshWs w;
    w.counts = ws;
    w.start = ws + cols;
    w.cursor = ws + 2 * cols + 1;
    w.part = ws + 3 * cols + 1;
    int64_t off = 3 * cols + 1 + scan_parts(cols);
    off = (off + 1) & ~(int64_t)1;  // 8-byte alignment
    w.total = reinterpret_cast<unsigned long long*>(ws + off);
    return w;
}

__device__ __forceinline__ float centre(int i, double dx) {
    return (float)(((double)i + 0.5) * dx);
}

// First index whose centre is >= lo (lo <= hi, both finite), clamped to [0, n].
__device__ int first_at_or_above(double lo, double dx, int n) {
    double q = ceil(lo / dx - 0.5);
    q = q < 0.0 ? 0.0 : (q > (double)n ? (double)n : q);
    return (int)q;
}

// Last index whose centre is <= hi, clamped to [-1, n - 1].
__device__ int last_at_or_below(double hi, double dx, int n) {
    double q = floor(hi / dx - 0.5);
    q = q < -1.0 ? -1.0 : (q > (double)(n - 1) ? (double)(n - 1) : q);
    return (int)q;
}

// Column range [j0, j1] x [k0, k1] of triangle t (empty when j0 > j1 or
// k0 > k1; also for non-finite vertices, which the test never counts).
// The box is widened by 1e-4 of the magnitudes involved (the domain and the
// triangle's coordinates): a ray passes the FP32 barycentric test only within
// a few ulps of those magnitudes of the triangle's y-z projection.
__device__ void tri_columns(const MeshArgs& a, int t, int& j0, int& j1, int& k0, int& k1) {
    const float* v = a.tris + 9 * (size_t)t;
    bool ok = true;
    double mag = a.extent;
    for (int e = 0; e < 9; ++e) {
        ok = ok && isfinite(v[e]);
        mag = fmax(mag, fabs((double)v[e]));
    }
    if (!ok) {
        j0 = 1; j1 = 0; k0 = 1; k1 = 0;
        return;
    }
    const double pad = 1e-4 * (1.0 + mag);
    const double ylo = fmin(fmin((double)v[1], (double)v[4]), (double)v[7]) - pad;
    const double yhi = fmax(fmax((double)v[1], (double)v[4]), (double)v[7]) + pad;
    const double zlo = fmin(fmin((double)v[2], (double)v[5]), (double)v[8]) - pad;
    const double zhi = fmax(fmax((double)v[2], (double)v[5]), (double)v[8]) + pad;
    j0 = first_at_or_above(ylo, a.dx, a.ny);
    j1 = last_at_or_below(yhi, a.dx, a.ny);
    k0 = first_at_or_above(zlo, a.dx, a.nz);
    k1 = last_at_or_below(zhi, a.dx, a.nz);
}

__global__ void k_mesh_clear(MeshWs w, int64_t cols) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < cols) {
        w.counts[c] = 0;
        w.cursor[c] = 0;
    }
    if (c == 0) *w.total = 0ull;
}

__global__ void k_mesh_count(MeshArgs a, MeshWs w) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tris) return;
    int j0, j1, k0, k1;
    tri_columns(a, t, j0, j1, k0, k1);
    if (j0 > j1 || k0 > k1) return;
    atomicAdd(w.total, (unsigned long long)(j1 - j0 + 1) * (unsigned long long)(k1 - k0 + 1));
    for (int j = j0; j <= j1; ++j)
        for (int k = k0; k <= k1; ++k) atomicAdd(&w.counts[(size_t)j * a.nz + k], 1);
}

__global__ void k_mesh_fill(MeshArgs a, MeshWs w, int32_t* __restrict__ list, int64_t n_entries) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_tris) return;
    int j0, j1, k0, k1;
    tri_columns(a, t, j0, j1, k0, k1);
    for (int j = j0; j <= j1; ++j)
        for (int k = k0; k <= k1; ++k) {
            const size_t col = (size_t)j * a.nz + k;
            const int64_t slot = (int64_t)w.start[col] + atomicAdd(&w.cursor[col], 1);
            if (slot < (int64_t)w.start[col + 1] && slot < n_entries) list[slot] = t;
        }
}

// Does the +x ray from p cross triangle v (mesh_marker.py:51-76)?
__device__ bool ray_crosses(const float* v, float px, float py, float pz) {
#pragma clang fp contract(off)
    const float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    const float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    const float hx = 0.0f, hy = -e2z, hz = e2y;
    const float det = e1x * hx + e1y * hy + e1z * hz;
    if (!(fabsf(det) > 1.0e-12f)) return false;
    const float f = 1.0f / det;
    const float sx = px - v[0], sy = py - v[1], sz = pz - v[2];
    const float u = f * (sx * hx + sy * hy + sz * hz);
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    const float qx = sy * e1z - sz * e1y;
    const float w = f * qx;
    if (!(w >= 0.0f && u + w <= 1.0f)) return false;
    const float qy = sz * e1x - sx * e1z;
    const float qz = sx * e1y - sy * e1x;
    const float t = f * (e2x * qx + e2y * qy + e2z * qz);
    return t > 1.0e-6f;
}

// One thread per cell (k fastest): crossing parity over the column's
// triangles; an inside cell that is not yet solid becomes solid and, when
// face arrays are given, writes the obstacle velocity into its six faces
// (k3_write_solid_face_vel, solver3d.py:816-836: two new neighbours write
// the same value to their shared face).
__global__ __launch_bounds__(GF_WG) void k_mesh_mark(MeshArgs a, const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ list,
                                                     int32_t* __restrict__ marker, float* su,
                                                     float* sv, float* sw, float vx, float vy,
                                                     float vz) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= a.nz || j >= a.ny) return;
    const size_t col = (size_t)j * a.nz + k;
    const float px = centre(i, a.dx), py = centre(j, a.dx), pz = centre(k, a.dx);
    int crossings = 0;
    for (int e = start[col], e1 = start[col + 1]; e < e1; ++e) {
        const int t = list[e];
        if ((unsigned)t < (unsigned)a.n_tris) crossings += ray_crosses(a.tris + 9 * (size_t)t, px, py, pz);
    }
    if (!(crossings & 1)) return;
    const size_t c = ((size_t)i * a.ny + j) * a.nz + k;
    if (marker[c] == 2) return;
    marker[c] = 2;
    if (!su) return;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    su[((size_t)i * ny + j) * nz + k] = vx;
    su[((size_t)(i + 1) * ny + j) * nz + k] = vx;
    sv[((size_t)i * (ny + 1) + j) * nz + k] = vy;
    sv[((size_t)i * (ny + 1) + j + 1) * nz + k] = vy;
    sw[((size_t)i * ny + j) * (nz + 1) + k] = vz;
    sw[((size_t)i * ny + j) * (nz + 1) + k + 1] = vz;
    (void)nx;
}

int mesh_check(const float* tris, int n_tris, int nx, int ny, int nz, double dx, const int32_t* ws,
               const char* who) {
    GF_REQUIRE(n_tris >= 0 && (n_tris == 0 || tris) && ws, "%s: null buffer", who);
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && dx > 0.0, "%s: grid", who);
    GF_REQUIRE((int64_t)ny * nz < INT32_MAX && (int64_t)nx * ny * nz < INT32_MAX,
               "%s: grid too large", who);
    GF_REQUIRE(nx <= 65535 && (ny + 3) / 4 <= 65535, "%s: grid %d x %d", who, nx, ny);
    return 0;
}

MeshArgs mesh_args(const float* tris, int n_tris, int nx, int ny, int nz, double dx) {
    MeshArgs a;
    a.tris = tris; a.n_tris = n_tris; a.nx = nx; a.ny = ny; a.nz = nz; a.dx = dx;
    a.extent = (double)(nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz)) * dx;
    return a;
}

inline dim3 grid_1d(int64_t n) {
    return dim3((unsigned)(n > 0 ? (n + GF_WG - 1) / GF_WG : 1));
}

}  // namespace

extern "C" int gf_mesh_workspace(int nx, int ny, int nz, int64_t* ws_ints) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && ws_ints, "gf_mesh_workspace: bad arguments");
    *ws_ints = mesh_ws_ints((int64_t)ny * nz);
    return 0;
}

extern "C" int gf_mesh_bin_sync(const float* tris, int n_tris, int nx, int ny, int nz, double dx,
                                int32_t* ws, int64_t* n_entries, void* stream) {
    int rc = mesh_check(tris, n_tris, nx, ny, nz, dx, ws, "gf_mesh_bin_sync");
    if (rc) return rc;
    GF_REQUIRE(n_entries, "gf_mesh_bin_sync: null output");
    const int64_t cols = (int64_t)ny * nz;
    const MeshArgs a = mesh_args(tris, n_tris, nx, ny, nz, dx);
    const MeshWs w = mesh_ws(ws, cols);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mesh_clear, grid_1d(cols), dim3(GF_WG), 0, st, w, cols);
    if (n_tris > 0)
        hipLaunchKernelGGL(k_mesh_count, grid_1d(n_tris), dim3(GF_WG), 0, st, a, w);
    GF_LAUNCH_CHECK("k_mesh_count");
    unsigned long long total = 0;
    hipError_t e = hipMemcpyAsync(&total, w.total, sizeof(total), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return gf::set_error((int)e, "gf_mesh_bin_sync: readback: %s", hipGetErrorString(e));
    GF_REQUIRE(total < (unsigned long long)INT32_MAX,
               "gf_mesh_bin_sync: %llu column entries exceed the int32 index range", total);
    rc = gf::scan_exclusive_i32(w.counts, (int)cols, w.start, w.part, st);
    if (rc) return rc;
    *n_entries = (int64_t)total;
    return 0;
}

extern "C" int gf_mesh_mark(const float* tris, int n_tris, int nx, int ny, int nz, double dx,
                            int32_t* ws, int32_t* list, int64_t n_entries, int32_t* marker,
                            float* solid_u, float* solid_v, float* solid_w, float vx, float vy,
                            float vz, void* stream) {
    int rc = mesh_check(tris, n_tris, nx, ny, nz, dx, ws, "gf_mesh_mark");
    if (rc) return rc;
    GF_REQUIRE(marker && n_entries >= 0 && (n_entries == 0 || list), "gf_mesh_mark: null buffer");
    GF_REQUIRE((solid_u && solid_v && solid_w) || (!solid_u && !solid_v && !solid_w),
               "gf_mesh_mark: give all three face arrays or none");
    const int64_t cols = (int64_t)ny * nz;
    const MeshArgs a = mesh_args(tris, n_tris, nx, ny, nz, dx);
    const MeshWs w = mesh_ws(ws, cols);
    hipStream_t st = (hipStream_t)stream;
    if (n_tris > 0 && n_entries > 0)
        hipLaunchKernelGGL(k_mesh_fill, grid_1d(n_tris), dim3(GF_WG), 0, st, a, w, list, n_entries);
    hipLaunchKernelGGL(k_mesh_mark, dim3((nz + 63) / 64, (ny + 3) / 4, nx), dim3(GF_WG), 0, st, a,
                       w.start, list, marker, solid_u, solid_v, solid_w, vx, vy, vz);
    GF_LAUNCH_CHECK("k_mesh_mark");
    return 0;
}
