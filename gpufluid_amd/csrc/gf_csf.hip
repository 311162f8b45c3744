// Surface tension by the continuum surface force (S2.14, solver3d.py:203-549,
// driver _apply_surface_tension 3230-3300): fluid indicator from the marker,
// boundary-aware 3x3x3 box blur passes, unit normal and curvature by central
// differences, the face impulse sigma*dt/rho * kappa_face * grad(chi), and the
// force balance that removes each component's mean impulse over the faces
// next to fluid.  Dense cell-centred passes over the whole grid, as the
// reference computes them; the balance sums are deterministic (per-workgroup
// FP64 partials reduced in a fixed order) where the reference adds FP32
// atomics in arrival order.  Products and sums are rounded one by one.
#include "gf_common.h"

namespace {

// Cell passes: 64 lanes along k x 4 along j per workgroup, one i per grid z.
struct Cell {
    int i, j, k;
    size_t c;
};

__device__ __forceinline__ bool cell_of(int nx, int ny, int nz, Cell& q) {
    q.k = blockIdx.x * 64 + (threadIdx.x & 63);
    q.j = blockIdx.y * 4 + (threadIdx.x >> 6);
    q.i = blockIdx.z;
    if (q.k >= nz || q.j >= ny || q.i >= nx) return false;
    q.c = ((size_t)q.i * ny + q.j) * nz + q.k;
    return true;
}

inline dim3 cell_grid(int nx, int ny, int nz) {
    return dim3((unsigned)((nz + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nx);
}

// k3_indicator_from_marker (solver3d.py:212-229)
__global__ __launch_bounds__(GF_WG) void k_csf_indicator(const int32_t* __restrict__ marker,
                                                         float* __restrict__ chi, int nx, int ny,
                                                         int nz) {
    Cell q;
    if (!cell_of(nx, ny, nz, q)) return;
    chi[q.c] = marker[q.c] == 1 ? 1.0f : 0.0f;
}

// k3_box_blur_centred (solver3d.py:235-256): mean over the in-range cells of
// the 3x3x3 block, summed with di outermost and dk innermost.
__global__ __launch_bounds__(GF_WG) void k_csf_blur(const float* __restrict__ src,
                                                    float* __restrict__ dst, int nx, int ny,
                                                    int nz) {
    Cell q;
    if (!cell_of(nx, ny, nz, q)) return;
    float s = 0.0f, n = 0.0f;
    for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj)
            for (int dk = -1; dk <= 1; ++dk) {
                const int ii = q.i + di, jj = q.j + dj, kk = q.k + dk;
                if (ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz) {
                    s += src[((size_t)ii * ny + jj) * nz + kk];
                    n += 1.0f;
                }
            }
    dst[q.c] = s / n;
}

// k3_csf_normal (solver3d.py:262-286): n = grad(chi) / (|grad(chi)| + 1e-8),
// zero on the outer cell layer.
__global__ __launch_bounds__(GF_WG) void k_csf_normal(const float* __restrict__ chi,
                                                      float* __restrict__ n_x,
                                                      float* __restrict__ n_y,
                                                      float* __restrict__ n_z, int nx, int ny,
                                                      int nz, float dx) {
#pragma clang fp contract(off)
    Cell q;
    if (!cell_of(nx, ny, nz, q)) return;
    if (q.i == 0 || q.j == 0 || q.k == 0 || q.i == nx - 1 || q.j == ny - 1 || q.k == nz - 1) {
        n_x[q.c] = 0.0f; n_y[q.c] = 0.0f; n_z[q.c] = 0.0f;
        return;
    }
    const size_t si = (size_t)ny * nz, sj = (size_t)nz;
    const float h = 2.0f * dx;
    const float gx = (chi[q.c + si] - chi[q.c - si]) / h;
    const float gy = (chi[q.c + sj] - chi[q.c - sj]) / h;
    const float gz = (chi[q.c + 1] - chi[q.c - 1]) / h;
    const float mag = sqrtf(gx * gx + gy * gy + gz * gz) + 1.0e-8f;
    n_x[q.c] = gx / mag;
    n_y[q.c] = gy / mag;
    n_z[q.c] = gz / mag;
}

// k3_csf_curvature (solver3d.py:292-313): kappa = -div(n), zero on the outer
// cell layer.
__global__ __launch_bounds__(GF_WG) void k_csf_curvature(const float* __restrict__ n_x,
                                                         const float* __restrict__ n_y,
                                                         const float* __restrict__ n_z,
                                                         float* __restrict__ kappa, int nx,
                                                         int ny, int nz, float dx) {
#pragma clang fp contract(off)
    Cell q;
    if (!cell_of(nx, ny, nz, q)) return;
    if (q.i == 0 || q.j == 0 || q.k == 0 || q.i == nx - 1 || q.j == ny - 1 || q.k == nz - 1) {
        kappa[q.c] = 0.0f;
        return;
    }
    const size_t si = (size_t)ny * nz, sj = (size_t)nz;
    const float h = 2.0f * dx;
    const float dnx = (n_x[q.c + si] - n_x[q.c - si]) / h;
    const float dny = (n_y[q.c + sj] - n_y[q.c - sj]) / h;
    const float dnz = (n_z[q.c + 1] - n_z[q.c - 1]) / h;
    kappa[q.c] = -(dnx + dny + dnz);
}

// Face passes: component c = blockIdx.z % 3, planes [chunk*32, chunk*32+32)
// along the face's own i axis, 64 lanes along k x 4 along j.
constexpr int CSF_CHUNK = 32;

struct FaceArgs {
    float* u; float* v; float* w;
    const float* chi; const float* kappa; const int32_t* marker;
    double* red;          // per workgroup: sum, count; then 3 means (NULL: not kept)
    int nx, ny, nz, gx, gy;
    float coeff, dx;
    int comp;             // >= 0: that component only (grid z = chunks), else all three
};

// Component and i-chunk of this workgroup.
__device__ __forceinline__ void face_part(const FaceArgs& a, int& c, int& chunk) {
    if (a.comp >= 0) {
        c = a.comp;
        chunk = blockIdx.z;
    } else {
        c = blockIdx.z % 3;
        chunk = blockIdx.z / 3;
    }
}

__host__ __device__ __forceinline__ int face_slots(const FaceArgs& a) {
    const int nchunk = (a.nx + 1 + CSF_CHUNK - 1) / CSF_CHUNK;
    return a.gx * a.gy * 3 * nchunk;
}

// The face between cells lo and hi along axis c; `active` = both cells in
// range and neither solid (the apply kernels' domain), `fluid` = one side
// fluid (the balance's domain).
struct Face {
    size_t f, lo, hi;
    bool active, fluid;
};

__device__ __forceinline__ Face face_at(const FaceArgs& a, int c, int i, int j, int k) {
    Face r;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int fx = nx + (c == 0), fy = ny + (c == 1), fz = nz + (c == 2);
    r.f = ((size_t)i * fy + j) * fz + k;
    const int ax = c == 0 ? i : (c == 1 ? j : k);
    const int n_ax = c == 0 ? nx : (c == 1 ? ny : nz);
    r.active = false;
    r.fluid = false;
    r.lo = r.hi = 0;
    if (ax <= 0 || ax >= n_ax) return r;
    r.hi = ((size_t)i * ny + j) * nz + k;
    r.lo = r.hi - (c == 0 ? (size_t)ny * nz : (c == 1 ? (size_t)nz : (size_t)1));
    const int mlo = a.marker[r.lo], mhi = a.marker[r.hi];
    r.active = mlo != 2 && mhi != 2;
    r.fluid = r.active && (mlo == 1 || mhi == 1);
    return r;
}

__device__ __forceinline__ void block_sum2(double& s, double& n, double* sh) {
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        n += __shfl_xor(n, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[2 * wave] = s; sh[2 * wave + 1] = n; }
    __syncthreads();
    if (threadIdx.x == 0) {
        s = sh[0]; n = sh[1];
        for (int w = 1; w < GF_WG / 64; ++w) { s += sh[2 * w]; n += sh[2 * w + 1]; }
    }
}

// k3_csf_apply_u/v/w (solver3d.py:316-385): face += coeff * kappa_face *
// grad(chi) on active faces; the workgroup also sums the change over faces
// next to fluid (k3_csf_sum_*, 400-464).
__global__ __launch_bounds__(GF_WG) void k_csf_apply(FaceArgs a) {
#pragma clang fp contract(off)
    __shared__ double sh[2 * GF_WG / 64];
    int c, chunk;
    face_part(a, c, chunk);
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int fx = a.nx + (c == 0), fy = a.ny + (c == 1), fz = a.nz + (c == 2);
    float* vel = c == 0 ? a.u : (c == 1 ? a.v : a.w);
    double s = 0.0, n = 0.0;
    if (k < fz && j < fy) {
        const int i1 = min((chunk + 1) * CSF_CHUNK, fx);
        for (int i = chunk * CSF_CHUNK; i < i1; ++i) {
            const Face q = face_at(a, c, i, j, k);
            if (!q.active) continue;
            const float grad = (a.chi[q.hi] - a.chi[q.lo]) / a.dx;
            const float kf = 0.5f * (a.kappa[q.hi] + a.kappa[q.lo]);
            const float before = vel[q.f];
            const float after = before + a.coeff * kf * grad;
            vel[q.f] = after;
            if (q.fluid) {
                s += (double)(after - before);
                n += 1.0;
            }
        }
    }
    if (!a.red) return;
    block_sum2(s, n, sh);
    if (threadIdx.x == 0) {
        const size_t slot = blockIdx.x + (size_t)a.gx * (blockIdx.y + (size_t)a.gy * blockIdx.z);
        a.red[2 * slot] = s;
        a.red[2 * slot + 1] = n;
    }
}

// k3_csf_sum_u/v/w (solver3d.py:400-464) for one component: per-workgroup
// FP64 partials of (after - before) and of the face count over the faces
// next to fluid; k_csf_sum_fin adds their fixed-order total, rounded to FP32,
// to sum[0] and count[0].
__global__ __launch_bounds__(GF_WG) void k_csf_sum(FaceArgs a, const float* __restrict__ before,
                                                   const float* __restrict__ after) {
#pragma clang fp contract(off)
    __shared__ double sh[2 * GF_WG / 64];
    const int c = a.comp, chunk = blockIdx.z;
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int fx = a.nx + (c == 0), fy = a.ny + (c == 1), fz = a.nz + (c == 2);
    double s = 0.0, n = 0.0;
    if (k < fz && j < fy) {
        const int i1 = min((chunk + 1) * CSF_CHUNK, fx);
        for (int i = chunk * CSF_CHUNK; i < i1; ++i) {
            const Face q = face_at(a, c, i, j, k);
            if (!q.fluid) continue;
            s += (double)(after[q.f] - before[q.f]);
            n += 1.0;
        }
    }
    block_sum2(s, n, sh);
    if (threadIdx.x == 0) {
        const size_t slot = blockIdx.x + (size_t)a.gx * (blockIdx.y + (size_t)a.gy * blockIdx.z);
        a.red[2 * slot] = s;
        a.red[2 * slot + 1] = n;
    }
}

__global__ __launch_bounds__(GF_WG) void k_csf_sum_fin(const double* red, int slots, float* sum,
                                                       float* count) {
    __shared__ double sh[2 * GF_WG / 64];
    double s = 0.0, n = 0.0;
    for (int e = threadIdx.x; e < slots; e += GF_WG) {
        s += red[2 * (size_t)e];
        n += red[2 * (size_t)e + 1];
    }
    block_sum2(s, n, sh);
    if (threadIdx.x == 0) {
        sum[0] = sum[0] + (float)s;
        count[0] = count[0] + (float)n;
    }
}

// k3_csf_subtract_bias_*_dev (solver3d.py:478-547) for one component:
// faces next to fluid lose sum[0] / count[0] when count[0] > 0.5.
__global__ __launch_bounds__(GF_WG) void k_csf_bias(FaceArgs a, const float* __restrict__ sum,
                                                    const float* __restrict__ count) {
    const int c = a.comp, chunk = blockIdx.z;
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int fx = a.nx + (c == 0), fy = a.ny + (c == 1), fz = a.nz + (c == 2);
    if (k >= fz || j >= fy) return;
    const float cnt = count[0];
    if (!(cnt > 0.5f)) return;
    const float mean = sum[0] / cnt;
    float* vel = c == 0 ? a.u : (c == 1 ? a.v : a.w);
    const int i1 = min((chunk + 1) * CSF_CHUNK, fx);
    for (int i = chunk * CSF_CHUNK; i < i1; ++i) {
        const Face q = face_at(a, c, i, j, k);
        if (q.fluid) vel[q.f] = vel[q.f] - mean;
    }
}

// Per component: sum the workgroups' partials in slot order, round to FP32
// like the reference's accumulators, and keep the FP32 mean (0 when no face
// counts, where the reference skips the subtraction).
__global__ __launch_bounds__(GF_WG) void k_csf_mean(double* red, int slots, int gz_per_c) {
    __shared__ double sh[2 * GF_WG / 64];
    const int c = blockIdx.x;
    const int per = slots / (3 * gz_per_c);  // slots per grid-z index
    double s = 0.0, n = 0.0;
    for (int z = c; z < 3 * gz_per_c; z += 3)
        for (int e = threadIdx.x; e < per; e += GF_WG) {
            const size_t slot = (size_t)z * per + e;
            s += red[2 * slot];
            n += red[2 * slot + 1];
        }
    block_sum2(s, n, sh);
    if (threadIdx.x == 0) {
        const float fs = (float)s, fn = (float)n;
        red[2 * (size_t)slots + c] = fn > 0.5f ? (double)(fs / fn) : 0.0;
    }
}

// k3_csf_subtract_bias_*_dev (solver3d.py:472-547)
__global__ __launch_bounds__(GF_WG) void k_csf_balance(FaceArgs a) {
    int c, chunk;
    face_part(a, c, chunk);
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int fx = a.nx + (c == 0), fy = a.ny + (c == 1), fz = a.nz + (c == 2);
    if (k >= fz || j >= fy) return;
    const float mean = (float)a.red[2 * (size_t)face_slots(a) + c];
    if (mean == 0.0f) return;
    float* vel = c == 0 ? a.u : (c == 1 ? a.v : a.w);
    const int i1 = min((chunk + 1) * CSF_CHUNK, fx);
    for (int i = chunk * CSF_CHUNK; i < i1; ++i) {
        const Face q = face_at(a, c, i, j, k);
        if (q.fluid) vel[q.f] = vel[q.f] - mean;
    }
}

}  // namespace

extern "C" int gf_csf_workspace(int nx, int ny, int nz, int64_t* work_floats,
                                int64_t* red_doubles) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && work_floats && red_doubles,
               "gf_csf_workspace: bad arguments");
    const int64_t gx = (nz + 1 + 63) / 64, gy = (ny + 1 + 3) / 4;
    const int64_t nchunk = (nx + 1 + CSF_CHUNK - 1) / CSF_CHUNK;
    *work_floats = 6LL * nx * ny * nz;
    *red_doubles = 2 * gx * gy * 3 * nchunk + 3;
    return 0;
}

extern "C" int gf_csf_apply(float* u, float* v, float* w, const int32_t* marker, int nx, int ny,
                            int nz, float dx, float coeff, int smoothing_passes, float* work,
                            double* red, void* stream) {
    GF_REQUIRE(nx > 2 && ny > 2 && nz > 2, "gf_csf_apply: grid %d x %d x %d", nx, ny, nz);
    GF_REQUIRE(nx + 1 <= 65535 && ny + 1 <= 4 * 65535, "gf_csf_apply: grid %d x %d x %d", nx,
               ny, nz);
    GF_REQUIRE(u && v && w && marker && work && red, "gf_csf_apply: null buffer");
    GF_REQUIRE(smoothing_passes >= 1 && smoothing_passes <= 64,
               "gf_csf_apply: smoothing_passes %d", smoothing_passes);
    GF_REQUIRE(dx > 0.0f, "gf_csf_apply: dx");
    hipStream_t st = (hipStream_t)stream;
    const size_t C = (size_t)nx * ny * nz;
    float* chi = work;
    float* tmp = work + C;
    float* n_x = work + 2 * C;
    float* n_y = work + 3 * C;
    float* n_z = work + 4 * C;
    float* kappa = work + 5 * C;
    const dim3 cg = cell_grid(nx, ny, nz);
    hipLaunchKernelGGL(k_csf_indicator, cg, dim3(GF_WG), 0, st, marker, chi, nx, ny, nz);
    for (int p = 0; p < smoothing_passes; ++p) {
        hipLaunchKernelGGL(k_csf_blur, cg, dim3(GF_WG), 0, st, chi, tmp, nx, ny, nz);
        float* t = chi; chi = tmp; tmp = t;
    }
    hipLaunchKernelGGL(k_csf_normal, cg, dim3(GF_WG), 0, st, chi, n_x, n_y, n_z, nx, ny, nz, dx);
    hipLaunchKernelGGL(k_csf_curvature, cg, dim3(GF_WG), 0, st, n_x, n_y, n_z, kappa, nx, ny, nz,
                       dx);
    FaceArgs a;
    a.u = u; a.v = v; a.w = w;
    a.chi = chi; a.kappa = kappa; a.marker = marker; a.red = red;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.gx = (nz + 1 + 63) / 64;
    a.gy = (ny + 1 + 3) / 4;
    a.coeff = coeff; a.dx = dx;
    a.comp = -1;
    const int nchunk = (nx + 1 + CSF_CHUNK - 1) / CSF_CHUNK;
    const dim3 fg((unsigned)a.gx, (unsigned)a.gy, (unsigned)(3 * nchunk));
    hipLaunchKernelGGL(k_csf_apply, fg, dim3(GF_WG), 0, st, a);
    hipLaunchKernelGGL(k_csf_mean, dim3(3), dim3(GF_WG), 0, st, red, face_slots(a), nchunk);
    hipLaunchKernelGGL(k_csf_balance, fg, dim3(GF_WG), 0, st, a);
    GF_LAUNCH_CHECK("surface tension");
    return 0;
}

// ---- the S2.14 sub-blocks one at a time (dense arrays, reference meaning) --

namespace {

int csf_grid_ok(int nx, int ny, int nz, const char* who) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "%s: grid %d x %d x %d", who, nx, ny, nz);
    GF_REQUIRE(nx + 1 <= 65535 && ny + 1 <= 4 * 65535, "%s: grid %d x %d x %d", who, nx, ny, nz);
    return 0;
}

FaceArgs one_comp(int comp, float* vel, const int32_t* marker, int nx, int ny, int nz) {
    FaceArgs a = {};
    a.u = comp == 0 ? vel : nullptr;
    a.v = comp == 1 ? vel : nullptr;
    a.w = comp == 2 ? vel : nullptr;
    a.marker = marker;
    a.nx = nx; a.ny = ny; a.nz = nz;
    a.gx = (nz + 1 + 63) / 64;
    a.gy = (ny + 1 + 3) / 4;
    a.comp = comp;
    return a;
}

dim3 comp_grid(const FaceArgs& a) {
    return dim3((unsigned)a.gx, (unsigned)a.gy, (unsigned)((a.nx + 1 + CSF_CHUNK - 1) / CSF_CHUNK));
}

}  // namespace

extern "C" int gf_k3_indicator_from_marker(const int32_t* marker, float* chi, int nx, int ny,
                                           int nz, void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_indicator_from_marker");
    if (rc) return rc;
    GF_REQUIRE(marker && chi, "gf_k3_indicator_from_marker: null buffer");
    hipLaunchKernelGGL(k_csf_indicator, cell_grid(nx, ny, nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, marker, chi, nx, ny, nz);
    GF_LAUNCH_CHECK("k_csf_indicator");
    return 0;
}

extern "C" int gf_k3_box_blur_centred(const float* src, float* dst, int nx, int ny, int nz,
                                      void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_box_blur_centred");
    if (rc) return rc;
    GF_REQUIRE(src && dst && src != dst, "gf_k3_box_blur_centred: bad buffers");
    hipLaunchKernelGGL(k_csf_blur, cell_grid(nx, ny, nz), dim3(GF_WG), 0, (hipStream_t)stream,
                       src, dst, nx, ny, nz);
    GF_LAUNCH_CHECK("k_csf_blur");
    return 0;
}

extern "C" int gf_k3_csf_normal(const float* chi, float* n_x, float* n_y, float* n_z, int nx,
                                int ny, int nz, float dx, void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_csf_normal");
    if (rc) return rc;
    GF_REQUIRE(chi && n_x && n_y && n_z, "gf_k3_csf_normal: null buffer");
    hipLaunchKernelGGL(k_csf_normal, cell_grid(nx, ny, nz), dim3(GF_WG), 0, (hipStream_t)stream,
                       chi, n_x, n_y, n_z, nx, ny, nz, dx);
    GF_LAUNCH_CHECK("k_csf_normal");
    return 0;
}

extern "C" int gf_k3_csf_curvature(const float* n_x, const float* n_y, const float* n_z,
                                   float* kappa, int nx, int ny, int nz, float dx, void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_csf_curvature");
    if (rc) return rc;
    GF_REQUIRE(n_x && n_y && n_z && kappa, "gf_k3_csf_curvature: null buffer");
    hipLaunchKernelGGL(k_csf_curvature, cell_grid(nx, ny, nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, n_x, n_y, n_z, kappa, nx, ny, nz, dx);
    GF_LAUNCH_CHECK("k_csf_curvature");
    return 0;
}

extern "C" int gf_k3_csf_apply(int comp, float* vel, const float* chi, const float* kappa,
                               const int32_t* marker, float coeff, float dx, int nx, int ny,
                               int nz, void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_csf_apply");
    if (rc) return rc;
    GF_REQUIRE(comp >= 0 && comp <= 2, "gf_k3_csf_apply: component %d", comp);
    GF_REQUIRE(vel && chi && kappa && marker, "gf_k3_csf_apply: null buffer");
    FaceArgs a = one_comp(comp, vel, marker, nx, ny, nz);
    a.chi = chi; a.kappa = kappa; a.coeff = coeff; a.dx = dx;
    hipLaunchKernelGGL(k_csf_apply, comp_grid(a), dim3(GF_WG), 0, (hipStream_t)stream, a);
    GF_LAUNCH_CHECK("k_csf_apply");
    return 0;
}

extern "C" int gf_k3_csf_sum_workspace(int nx, int ny, int nz, int64_t* red_doubles) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && red_doubles, "gf_k3_csf_sum_workspace: bad arguments");
    const int64_t gx = (nz + 1 + 63) / 64, gy = (ny + 1 + 3) / 4;
    *red_doubles = 2 * gx * gy * ((nx + 1 + CSF_CHUNK - 1) / CSF_CHUNK);
    return 0;
}

extern "C" int gf_k3_csf_sum(int comp, const float* before, const float* after,
                             const int32_t* marker, float* sum, float* count, int nx, int ny,
                             int nz, double* red, void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_csf_sum");
    if (rc) return rc;
    GF_REQUIRE(comp >= 0 && comp <= 2, "gf_k3_csf_sum: component %d", comp);
    GF_REQUIRE(before && after && marker && sum && count && red, "gf_k3_csf_sum: null buffer");
    FaceArgs a = one_comp(comp, nullptr, marker, nx, ny, nz);
    a.red = red;
    const dim3 g = comp_grid(a);
    hipLaunchKernelGGL(k_csf_sum, g, dim3(GF_WG), 0, (hipStream_t)stream, a, before, after);
    hipLaunchKernelGGL(k_csf_sum_fin, dim3(1), dim3(GF_WG), 0, (hipStream_t)stream, red,
                       (int)(g.x * g.y * g.z), sum, count);
    GF_LAUNCH_CHECK("k_csf_sum");
    return 0;
}

extern "C" int gf_k3_csf_subtract_bias(int comp, float* vel, const int32_t* marker,
                                       const float* sum, const float* count, int nx, int ny,
                                       int nz, void* stream) {
    int rc = csf_grid_ok(nx, ny, nz, "gf_k3_csf_subtract_bias");
    if (rc) return rc;
    GF_REQUIRE(comp >= 0 && comp <= 2, "gf_k3_csf_subtract_bias: component %d", comp);
    GF_REQUIRE(vel && marker && sum && count, "gf_k3_csf_subtract_bias: null buffer");
    FaceArgs a = one_comp(comp, vel, marker, nx, ny, nz);
    hipLaunchKernelGGL(k_csf_bias, comp_grid(a), dim3(GF_WG), 0, (hipStream_t)stream, a, sum,
                       count);
    GF_LAUNCH_CHECK("k_csf_bias");
    return 0;
}
