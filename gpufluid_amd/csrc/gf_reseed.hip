// Particle-count-changing passes between frames: the GPU reseed (S2.11.GPU,
// sim/reseed.py:161-396) and outflow stream compaction (D4.7.GPU,
// solver3d.py:1947-1985, 2659-2730).  Arrays are in the caller's particle
// order.  Every count and offset comes from integer atomics and scans, and
// the particles a reseed keeps are the first max_per_cell of each cell in
// index order, so the results are deterministic (the reference keeps the
// first arrivals of its atomics).
#include "gf_common.h"

namespace {

// Workspace layout (int32) for n particles and C cells.
struct ReseedWs {
    int32_t *keys, *tmp, *alive, *prefix, *counts, *cell_start, *deficit, *emit_off, *part;
};

inline int64_t scan_parts(int64_t n) { return (n + 4095) / 4096 + 1; }

ReseedWs reseed_ws(int32_t* ws, int64_t n, int64_t C) {
    ReseedWs w;
    int64_t o = 0;
    w.keys = ws + o; o += n;
    w.tmp = ws + o; o += n;
    w.alive = ws + o; o += n;
    w.prefix = ws + o; o += n + 1;
    w.counts = ws + o; o += C;
    w.cell_start = ws + o; o += C + 1;
    w.deficit = ws + o; o += C;
    w.emit_off = ws + o; o += C + 1;
    w.part = ws + o;
    return w;
}

int64_t reseed_ws_ints(int64_t n, int64_t C) {
    return 4 * n + 1 + 4 * C + 2 + std::max(scan_parts(n), scan_parts(C + 1));
}

__global__ void k_fill0(int32_t* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

inline dim3 grid_of(int64_t n) { return dim3((unsigned)std::max<int64_t>((n + GF_WG - 1) / GF_WG, 1)); }

// k_count_particles_per_cell (reseed.py:160-179): cell = clamp(int(p/dx)).
__global__ void k_rs_bin(const float* __restrict__ pos, int n, float dx, int nx, int ny, int nz,
                         int32_t* __restrict__ keys, int32_t* __restrict__ counts) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = gf::bin_axis(pos[3 * p + 0], dx, nx);
    const int j = gf::bin_axis(pos[3 * p + 1], dx, ny);
    const int k = gf::bin_axis(pos[3 * p + 2], dx, nz);
    const int key = (i * ny + j) * nz + k;
    keys[p] = key;
    atomicAdd(&counts[key], 1);
}

// Emission count per cell (reseed.py:336-343): cells that are fluid in the
// marker or hold particles, and are not solid, get min_per_cell - count.
__global__ void k_rs_deficit(const int32_t* __restrict__ marker,
                             const int32_t* __restrict__ counts, int64_t C, int min_per_cell,
                             int32_t* __restrict__ deficit) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int m = marker[c], cnt = counts[c];
    const bool fluid = (m == 1 || cnt > 0) && m != 2;
    deficit[c] = fluid ? max(min_per_cell - cnt, 0) : 0;
}

// Unstable slot of each particle inside its cell's range (consumes counts).
__global__ void k_rs_place(const int32_t* __restrict__ keys, int n,
                           int32_t* __restrict__ counts, const int32_t* __restrict__ cell_start,
                           int32_t* __restrict__ tmp) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int key = keys[p];
    const int old = atomicSub(&counts[key], 1);
    tmp[cell_start[key] + old - 1] = p;
}

// Rank among the cell's members by index (k_mark_keep_by_rank, reseed.py:185-215,
// with the arrival order replaced by index order): keep the first
// max_per_cell; particles in solid cells are dropped when drop_solid, else
// kept and not counted against the cell (the host reseed, reseed.py:112-136).
__global__ void k_rs_alive(const int32_t* __restrict__ keys, int n,
                           const int32_t* __restrict__ cell_start,
                           const int32_t* __restrict__ tmp, const int32_t* __restrict__ marker,
                           int max_per_cell, int drop_solid, int32_t* __restrict__ alive) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int key = keys[p];
    if (marker[key] == 2) {
        alive[p] = drop_solid ? 0 : 1;
        return;
    }
    const int s = cell_start[key], e = cell_start[key + 1];
    int rank = 0;
    for (int q = s; q < e; ++q) rank += tmp[q] < p;
    alive[p] = rank < max_per_cell ? 1 : 0;
}

// k3_mark_alive_outflows (solver3d.py:1946-1963)
__global__ void k_outflow_mark(const float* __restrict__ pos, int n,
                               const float* __restrict__ boxes, int n_box,
                               int32_t* __restrict__ alive) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float x = pos[3 * p + 0], y = pos[3 * p + 1], z = pos[3 * p + 2];
    int a = 1;
    for (int b = 0; b < n_box; ++b) {
        const float* q = boxes + 6 * b;
        if (x >= q[0] && x <= q[3] && y >= q[1] && y <= q[4] && z >= q[2] && z <= q[5]) a = 0;
    }
    alive[p] = a;
}

// k3_scatter_alive / k_scatter_alive_* (solver3d.py:1966-1982,
// reseed.py:218-264): row i moves to row prefix[i] when alive.
__global__ void k_compact_rows(const float* __restrict__ in, int width, int n,
                               const int32_t* __restrict__ alive,
                               const int32_t* __restrict__ prefix, float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * width) return;
    const int i = (int)(t / width), c = (int)(t - (int64_t)i * width);
    if (alive[i]) out[(int64_t)prefix[i] * width + c] = in[t];
}

// Emitted positions (reseed.py:345-352): each cell with a deficit d gets d
// rows from its offset on, at (cell + 0.5 + jitter) * dx.
__global__ void k_rs_emit(const int32_t* __restrict__ deficit,
                          const int32_t* __restrict__ emit_off, int nx, int ny, int nz, float dx,
                          const float* __restrict__ jitter, float* __restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)nx * ny * nz) return;
    const int d = deficit[c];
    if (d <= 0) return;
    const int k = (int)(c % nz), j = (int)((c / nz) % ny), i = (int)(c / ((int64_t)nz * ny));
    const float cf[3] = {(float)i, (float)j, (float)k};
    const int o = emit_off[c];
    for (int e = 0; e < d; ++e) {
        const int64_t r = (int64_t)(o + e) * 3;
        for (int a = 0; a < 3; ++a) out[r + a] = (cf[a] + 0.5f + jitter[r + a]) * dx;
    }
}

}  // namespace

extern "C" int gf_reseed_workspace(int n, int nx, int ny, int nz, int64_t* ws_ints) {
    GF_REQUIRE(n >= 0 && nx > 0 && ny > 0 && nz > 0 && ws_ints, "gf_reseed_workspace: bad arguments");
    *ws_ints = reseed_ws_ints(n, (int64_t)nx * ny * nz);
    return 0;
}

extern "C" int gf_reseed_plan_sync(const float* pos, int n, const int32_t* marker, int nx, int ny,
                                   int nz, float dx, int min_per_cell, int max_per_cell,
                                   int drop_solid, int32_t* ws, int* counts_host,
                                   void* stream) {
    GF_REQUIRE(n >= 0 && nx > 0 && ny > 0 && nz > 0 && dx > 0.0f, "gf_reseed_plan_sync: sizes");
    GF_REQUIRE((int64_t)nx * ny * nz < INT32_MAX, "gf_reseed_plan_sync: grid too large");
    GF_REQUIRE(marker && ws && counts_host && (n == 0 || pos), "gf_reseed_plan_sync: null buffer");
    GF_REQUIRE(max_per_cell >= 0, "gf_reseed_plan_sync: max_per_cell %d", max_per_cell);
    hipStream_t st = (hipStream_t)stream;
    const int64_t C = (int64_t)nx * ny * nz;
    ReseedWs w = reseed_ws(ws, n, C);
    hipLaunchKernelGGL(k_fill0, grid_of(C), dim3(GF_WG), 0, st, w.counts, C);
    if (n > 0)
        hipLaunchKernelGGL(k_rs_bin, grid_of(n), dim3(GF_WG), 0, st, pos, n, dx, nx, ny, nz,
                           w.keys, w.counts);
    int rc = gf::scan_exclusive_i32(w.counts, (int)C, w.cell_start, w.part, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rs_deficit, grid_of(C), dim3(GF_WG), 0, st, marker, w.counts, C,
                       min_per_cell, w.deficit);
    rc = gf::scan_exclusive_i32(w.deficit, (int)C, w.emit_off, w.part, st);
    if (rc) return rc;
    if (n > 0) {
        hipLaunchKernelGGL(k_rs_place, grid_of(n), dim3(GF_WG), 0, st, w.keys, n, w.counts,
                           w.cell_start, w.tmp);
        hipLaunchKernelGGL(k_rs_alive, grid_of(n), dim3(GF_WG), 0, st, w.keys, n, w.cell_start,
                           w.tmp, marker, max_per_cell, drop_solid, w.alive);
    }
    rc = gf::scan_exclusive_i32(w.alive, n, w.prefix, w.part, st);
    if (rc) return rc;
    GF_LAUNCH_CHECK("reseed plan");
    int32_t h[2] = {0, 0};
    if (hipMemcpyAsync(&h[0], w.prefix + n, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&h[1], w.emit_off + C, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return gf::set_error(GF_EINVAL, "gf_reseed_plan_sync: readback failed");
    if (n == 0) h[0] = 0;
    counts_host[0] = h[0];
    counts_host[1] = h[1];
    return 0;
}

extern "C" int gf_reseed_compact(const float* in, int width, int n, int nx, int ny, int nz,
                                 const int32_t* ws, float* out, void* stream) {
    GF_REQUIRE(width >= 1 && width <= 16 && n >= 0, "gf_reseed_compact: width %d", width);
    if (n == 0) return 0;
    GF_REQUIRE(in && out && ws && in != out, "gf_reseed_compact: bad buffers");
    ReseedWs w = reseed_ws(const_cast<int32_t*>(ws), n, (int64_t)nx * ny * nz);
    hipLaunchKernelGGL(k_compact_rows, grid_of((int64_t)n * width), dim3(GF_WG), 0,
                       (hipStream_t)stream, in, width, n, w.alive, w.prefix, out);
    GF_LAUNCH_CHECK("k_compact_rows");
    return 0;
}

extern "C" int gf_reseed_emit(int n, int nx, int ny, int nz, float dx, const int32_t* ws,
                              const float* jitter, float* out, void* stream) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && ws && jitter && out, "gf_reseed_emit: bad arguments");
    const int64_t C = (int64_t)nx * ny * nz;
    ReseedWs w = reseed_ws(const_cast<int32_t*>(ws), n, C);
    hipLaunchKernelGGL(k_rs_emit, grid_of(C), dim3(GF_WG), 0, (hipStream_t)stream, w.deficit,
                       w.emit_off, nx, ny, nz, dx, jitter, out);
    GF_LAUNCH_CHECK("k_rs_emit");
    return 0;
}

extern "C" int gf_outflow_mark_sync(const float* pos, int n, const float* boxes, int n_box,
                                    int32_t* alive, int32_t* prefix, int32_t* part,
                                    int* n_alive_host, void* stream) {
    GF_REQUIRE(n >= 0 && n_box >= 0 && n_alive_host, "gf_outflow_mark_sync: sizes");
    GF_REQUIRE(n == 0 || (pos && alive && prefix && part), "gf_outflow_mark_sync: null buffer");
    GF_REQUIRE(n_box == 0 || boxes, "gf_outflow_mark_sync: null boxes");
    hipStream_t st = (hipStream_t)stream;
    if (n == 0) { *n_alive_host = 0; return 0; }
    hipLaunchKernelGGL(k_outflow_mark, grid_of(n), dim3(GF_WG), 0, st, pos, n, boxes, n_box,
                       alive);
    int rc = gf::scan_exclusive_i32(alive, n, prefix, part, st);
    if (rc) return rc;
    GF_LAUNCH_CHECK("k_outflow_mark");
    int32_t h = 0;
    if (hipMemcpyAsync(&h, prefix + n, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return gf::set_error(GF_EINVAL, "gf_outflow_mark_sync: readback failed");
    *n_alive_host = h;
    return 0;
}

extern "C" int gf_compact_rows(const float* in, int width, int n, const int32_t* alive,
                               const int32_t* prefix, float* out, void* stream) {
    GF_REQUIRE(width >= 1 && width <= 16 && n >= 0, "gf_compact_rows: width %d", width);
    if (n == 0) return 0;
    GF_REQUIRE(in && out && alive && prefix && in != out, "gf_compact_rows: bad buffers");
    hipLaunchKernelGGL(k_compact_rows, grid_of((int64_t)n * width), dim3(GF_WG), 0,
                       (hipStream_t)stream, in, width, n, alive, prefix, out);
    GF_LAUNCH_CHECK("k_compact_rows");
    return 0;
}
