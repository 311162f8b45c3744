// Adjoint of one FLIP/PIC substep (the reverse of gf_scene_sort_sync ->
// gf_scene_p2g -> pressure -> gf_scene_g2p), for differentiable rollouts.
//
// Forward map of a substep (reference arithmetic, solver3d.py:72-156,
// 758-799, 846-871, 1364-1463; gridmath.py:30-95), per face f and particle p:
//   m_f = sum_p w_pf(x_p),  q_f = sum_p w_pf v_p,  u_n = q/m (m > 1e-8, else 0)
//   u_pre = Mb (u_n + g dt e_y)          (Mb: 0 on domain-edge / solid faces)
//   b = s_d D u_pre on fluid cells       (s_d = rho dx / dt)
//   A p = sign * b                        (pressure solve, A SPD on fluid cells)
//   u_new = Mb (u_pre - s_g E G p)       (s_g = dt / (rho dx); E: eligible faces)
//   v'_p = a v_p + S(u_new, x_p) - a S(u_n, x_p)      (FLIP blend a; PIC: a = 0)
//   x'_p = clamp(x_p + v'_p dt, 1.001 dx, dom - 1.001 dx)
// The reverse pass, given the adjoints of (x', v'):
//   gf_adj_g2p:      clamp mask, v-bar, x-bar from the sample gradients, and the
//                    transposed gather (a scatter) into u_new-bar and u_n-bar
//   gf_adj_pressure_rhs: p-bar = -s_g G^T E u_new-bar on fluid cells (compact)
//   (second PCG on the same matrix: b-bar = sign * A^-1 p-bar)
//   gf_adj_faces:    u_n-bar = Mb (Mb u_new-bar + s_d D^T b-bar) + u_n-bar(G2P);
//                    q-bar = u_n-bar / m, m-bar = -u_n-bar u_n / m
//   gf_adj_p2g:      the transposed scatter (a gather): v-bar += sum w q-bar,
//                    x-bar += sum grad(w) (q-bar v + m-bar)
// The fluid/solid masks and the binning are piecewise constant in x and are
// treated as constants.  Particle adjoints are indexed by particle id (pid),
// so the per-substep sort permutation needs no separate reverse step.
#include "gf_common.h"

namespace {

struct AdjDims {
    int nx, ny, nz;
    float dx;
};

__device__ __forceinline__ size_t face_index(int comp, int i, int j, int k, const AdjDims& d) {
    if (comp == 0) return ((size_t)i * d.ny + j) * d.nz + k;
    if (comp == 1) return ((size_t)i * (d.ny + 1) + j) * d.nz + k;
    return ((size_t)i * d.ny + j) * (d.nz + 1) + k;
}

__device__ __forceinline__ void face_extent(int comp, const AdjDims& d, int& ex, int& ey, int& ez) {
    ex = d.nx + (comp == 0);
    ey = d.ny + (comp == 1);
    ez = d.nz + (comp == 2);
}

__global__ void k_adj_fill(float* a, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

// Face-space coordinates of particle (lx,ly,lz) = pos/dx for component comp.
__device__ __forceinline__ void face_coords(int comp, float lx, float ly, float lz, float& cx,
                                            float& cy, float& cz) {
    cx = comp == 0 ? lx : lx - 0.5f;
    cy = comp == 1 ? ly : ly - 0.5f;
    cz = comp == 2 ? lz : lz - 0.5f;
}

// ---- P2G weights (gridmath.py:71-95): corners floor(c), floor(c)+1; corners
// outside the face array are dropped.
__global__ void k_adj_weights(const float* __restrict__ pos, int n, AdjDims d, float* mu,
                              float* mv, float* mw) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float lx = pos[3 * p] / d.dx, ly = pos[3 * p + 1] / d.dx, lz = pos[3 * p + 2] / d.dx;
    float* m[3] = {mu, mv, mw};
    for (int comp = 0; comp < 3; ++comp) {
        float cx, cy, cz;
        face_coords(comp, lx, ly, lz, cx, cy, cz);
        const int i0 = (int)floorf(cx), j0 = (int)floorf(cy), k0 = (int)floorf(cz);
        const float fx = cx - (float)i0, fy = cy - (float)j0, fz = cz - (float)k0;
        int ex, ey, ez;
        face_extent(comp, d, ex, ey, ez);
        for (int c = 0; c < 8; ++c) {
            const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
            const int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
            if (ii < 0 || ii >= ex || jj < 0 || jj >= ey || kk < 0 || kk >= ez) continue;
            const float wgt = (di ? fx : 1.0f - fx) * (dj ? fy : 1.0f - fy) * (dk ? fz : 1.0f - fz);
            atomicAdd(&m[comp][face_index(comp, ii, jj, kk, d)], wgt);
        }
    }
}

// ---- sample3 (gridmath.py:30-65) with its gradient and corner weights.
struct Sample {
    size_t idx[8];
    float wgt[8];
    float gx, gy, gz;   // d sample / d (lx, ly, lz) of field f
    float val;
};

__device__ __forceinline__ float sample_grad(const float* __restrict__ f, int comp,
                                             float fx_raw, float fy_raw, float fz_raw,
                                             const AdjDims& d, Sample& s) {
    int ex, ey, ez;
    face_extent(comp, d, ex, ey, ez);
    const float fx = fminf(fmaxf(fx_raw, 0.0f), (float)ex);
    const float fy = fminf(fmaxf(fy_raw, 0.0f), (float)ey);
    const float fz = fminf(fmaxf(fz_raw, 0.0f), (float)ez);
    const int i0 = gf::clampi((int)floorf(fx), 0, ex - 1);
    const int j0 = gf::clampi((int)floorf(fy), 0, ey - 1);
    const int k0 = gf::clampi((int)floorf(fz), 0, ez - 1);
    const int i1 = min(i0 + 1, ex - 1), j1 = min(j0 + 1, ey - 1), k1 = min(k0 + 1, ez - 1);
    const float rx = fx - (float)i0, ry = fy - (float)j0, rz = fz - (float)k0;
    const float sx = fminf(fmaxf(rx, 0.0f), 1.0f);
    const float sy = fminf(fmaxf(ry, 0.0f), 1.0f);
    const float sz = fminf(fmaxf(rz, 0.0f), 1.0f);
    // d sx / d fx_raw is 1 only where neither clamp is active
    const float dx_ok = (fx_raw >= 0.0f && fx_raw <= (float)ex && rx >= 0.0f && rx <= 1.0f) ? 1.0f : 0.0f;
    const float dy_ok = (fy_raw >= 0.0f && fy_raw <= (float)ey && ry >= 0.0f && ry <= 1.0f) ? 1.0f : 0.0f;
    const float dz_ok = (fz_raw >= 0.0f && fz_raw <= (float)ez && rz >= 0.0f && rz <= 1.0f) ? 1.0f : 0.0f;
    const int is[2] = {i0, i1}, js[2] = {j0, j1}, ks[2] = {k0, k1};
    float v[8];
    for (int c = 0; c < 8; ++c) {
        const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
        s.idx[c] = face_index(comp, is[di], js[dj], ks[dk], d);
        s.wgt[c] = (di ? sx : 1.0f - sx) * (dj ? sy : 1.0f - sy) * (dk ? sz : 1.0f - sz);
        v[c] = f ? f[s.idx[c]] : 0.0f;
    }
    float val = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for (int c = 0; c < 8; ++c) {
        const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
        val += s.wgt[c] * v[c];
        const float wx = di ? sx : 1.0f - sx, wy = dj ? sy : 1.0f - sy, wz = dk ? sz : 1.0f - sz;
        const float sgx = di ? 1.0f : -1.0f, sgy = dj ? 1.0f : -1.0f, sgz = dk ? 1.0f : -1.0f;
        gx += sgx * wy * wz * v[c];
        gy += wx * sgy * wz * v[c];
        gz += wx * wy * sgz * v[c];
    }
    s.gx = gx * dx_ok;
    s.gy = gy * dy_ok;
    s.gz = gz * dz_ok;
    s.val = val;
    return val;
}

struct G2PAdj {
    const float* pos; const float* vel; const int32_t* pid;
    const float* f[3];     // u_new per component
    const float* fs[3];    // u_n (us) per component
    float* af[3];          // adjoint of u_new (accumulated)
    float* afs[3];         // adjoint of u_n (accumulated)
    float* apos; float* avel;
    int n;
    AdjDims d;
    float fb, dt, dom[3];
};

// Reverse of FLIP/PIC G2P + advection + clamp (solver3d.py:1414-1463) for one
// particle; apos/avel hold the adjoints of (x', v') on entry and of (x, v) on exit.
__global__ __launch_bounds__(GF_WG) void k_adj_g2p(G2PAdj a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int id = a.pid[p];
    const float inv = 1.0f / a.d.dx;
    const float x[3] = {a.pos[3 * p], a.pos[3 * p + 1], a.pos[3 * p + 2]};
    const float lx = x[0] / a.d.dx, ly = x[1] / a.d.dx, lz = x[2] / a.d.dx;
    Sample sn[3], so[3];
    float vnew[3];
    for (int c = 0; c < 3; ++c) {
        float cx, cy, cz;
        face_coords(c, lx, ly, lz, cx, cy, cz);
        const float pic = sample_grad(a.f[c], c, cx, cy, cz, a.d, sn[c]);
        const float old = sample_grad(a.fs[c], c, cx, cy, cz, a.d, so[c]);
        const float flip = a.vel[3 * p + c] + (pic - old);
        vnew[c] = a.fb * flip + (1.0f - a.fb) * pic;
    }
    const float eps = a.d.dx * 1.001f;
    float xb[3], vb[3];
    for (int c = 0; c < 3; ++c) {
        const float xp = x[c] + vnew[c] * a.dt;
        const float inside = (xp > eps && xp < a.dom[c] - eps) ? 1.0f : 0.0f;
        xb[c] = a.apos[3 * id + c] * inside;          // adjoint of the unclamped position
        vb[c] = a.avel[3 * id + c] + a.dt * xb[c];    // adjoint of v'
    }
    float gx = xb[0], gy = xb[1], gz = xb[2];
    for (int c = 0; c < 3; ++c) {
        gx += vb[c] * (sn[c].gx - a.fb * so[c].gx) * inv;
        gy += vb[c] * (sn[c].gy - a.fb * so[c].gy) * inv;
        gz += vb[c] * (sn[c].gz - a.fb * so[c].gz) * inv;
        for (int k = 0; k < 8; ++k) {
            atomicAdd(&a.af[c][sn[c].idx[k]], sn[c].wgt[k] * vb[c]);
            atomicAdd(&a.afs[c][so[c].idx[k]], -a.fb * so[c].wgt[k] * vb[c]);
        }
    }
    a.apos[3 * id] = gx;
    a.apos[3 * id + 1] = gy;
    a.apos[3 * id + 2] = gz;
    for (int c = 0; c < 3; ++c) a.avel[3 * id + c] = a.fb * vb[c];
}

__device__ __forceinline__ int cell_marker(const int32_t* m, int i, int j, int k, const AdjDims& d) {
    if (i < 0 || j < 0 || k < 0 || i >= d.nx || j >= d.ny || k >= d.nz) return -1;
    return m[((size_t)i * d.ny + j) * d.nz + k];
}

__device__ __forceinline__ long long compact_index(const int32_t* tile_map, int i, int j, int k,
                                                   const AdjDims& d) {
    const int tny = (d.ny + 7) >> 3, tnz = (d.nz + 7) >> 3;
    const int slot = tile_map[((size_t)(i >> 3) * tny + (j >> 3)) * tnz + (k >> 3)];
    if (slot < 0) return -1;
    return (long long)slot * GF_TILE_CELLS + (((i & 7) * 8 + (j & 7)) * 8 + (k & 7));
}

// Mask of faces the solid BC leaves free (not on the domain edge, no solid
// neighbour), solver3d.py:758-799.
__device__ __forceinline__ bool face_free(int comp, int i, int j, int k, const int32_t* m,
                                          const AdjDims& d) {
    int n = comp == 0 ? d.nx : (comp == 1 ? d.ny : d.nz);
    int t = comp == 0 ? i : (comp == 1 ? j : k);
    if (t == 0 || t == n) return false;
    const int ml = cell_marker(m, i - (comp == 0), j - (comp == 1), k - (comp == 2), d);
    const int mr = cell_marker(m, i, j, k, d);
    return ml != GF_SOLID && mr != GF_SOLID;
}

// p-bar on fluid cells: -s_g G^T E u_new-bar, written to the compact vector
// (every entry of an active tile is written; non-fluid entries get 0).
__global__ void k_adj_pressure_rhs(const int32_t* __restrict__ marker,
                                   const int32_t* __restrict__ tile_map, const float* au,
                                   const float* av, const float* aw, AdjDims d, float sg,
                                   float* bc) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t C = (size_t)d.nx * d.ny * d.nz;
    if (cell >= C) return;
    const int k = (int)(cell % d.nz), j = (int)((cell / d.nz) % d.ny), i = (int)(cell / ((size_t)d.ny * d.nz));
    const long long ci = compact_index(tile_map, i, j, k, d);
    if (ci < 0) return;
    float pb = 0.0f;
    if (marker[cell] == GF_FLUID) {
        const float* af[3] = {au, av, aw};
        for (int comp = 0; comp < 3; ++comp) {
            const int oi = comp == 0, oj = comp == 1, ok = comp == 2;
            // upper face (this cell is the lower side l): +s_g
            const int mu = cell_marker(marker, i + oi, j + oj, k + ok, d);
            if (mu >= 0 && mu != GF_SOLID)
                pb += sg * af[comp][face_index(comp, i + oi, j + oj, k + ok, d)];
            // lower face (this cell is the upper side r): -s_g
            const int ml = cell_marker(marker, i - oi, j - oj, k - ok, d);
            if (ml >= 0 && ml != GF_SOLID)
                pb -= sg * af[comp][face_index(comp, i, j, k, d)];
        }
    }
    bc[ci] = pb;
}

// u_n-bar = Mb (Mb u_new-bar + s_d D^T b-bar) + (u_n-bar from G2P), then
// q-bar = u_n-bar/m, m-bar = -u_n-bar u_n/m (m > 1e-8), written over af / afs.
template <int COMP>
__global__ void k_adj_faces(const int32_t* __restrict__ marker,
                            const int32_t* __restrict__ tile_map, const float* __restrict__ bbar,
                            const float* __restrict__ m, const float* __restrict__ un, float* af,
                            float* afs, AdjDims d, float sd) {
    int ex, ey, ez;
    face_extent(COMP, d, ex, ey, ez);
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (size_t)ex * ey * ez) return;
    const int k = (int)(f % ez), j = (int)((f / ez) % ey), i = (int)(f / ((size_t)ey * ez));
    const bool freef = face_free(COMP, i, j, k, marker, d);
    float g = freef ? af[f] : 0.0f;
    // divergence transpose: +s_d b-bar of the lower cell, -s_d b-bar of the upper cell
    const int li = i - (COMP == 0), lj = j - (COMP == 1), lk = k - (COMP == 2);
    if (cell_marker(marker, li, lj, lk, d) == GF_FLUID) {
        const long long c = compact_index(tile_map, li, lj, lk, d);
        if (c >= 0) g += sd * bbar[c];
    }
    if (cell_marker(marker, i, j, k, d) == GF_FLUID) {
        const long long c = compact_index(tile_map, i, j, k, d);
        if (c >= 0) g -= sd * bbar[c];
    }
    g = freef ? g : 0.0f;
    const float ubar = g + afs[f];
    const float mf = m[f];
    float qb = 0.0f, mb = 0.0f;
    if (mf > 1e-8f) {
        qb = ubar / mf;
        mb = -ubar * un[f] / mf;
    }
    af[f] = qb;
    afs[f] = mb;
}

struct P2GAdj {
    const float* pos; const float* vel; const int32_t* pid;
    const float* qb[3]; const float* mb[3];
    float* apos; float* avel;
    int n;
    AdjDims d;
};

// Transposed P2G scatter: a gather of q-bar and m-bar over the particle's
// corners; adds to the particle adjoints.
__global__ __launch_bounds__(GF_WG) void k_adj_p2g(P2GAdj a) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const int id = a.pid[p];
    const float inv = 1.0f / a.d.dx;
    const float lx = a.pos[3 * p] / a.d.dx, ly = a.pos[3 * p + 1] / a.d.dx,
                lz = a.pos[3 * p + 2] / a.d.dx;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    for (int comp = 0; comp < 3; ++comp) {
        float cx, cy, cz;
        face_coords(comp, lx, ly, lz, cx, cy, cz);
        const int i0 = (int)floorf(cx), j0 = (int)floorf(cy), k0 = (int)floorf(cz);
        const float fx = cx - (float)i0, fy = cy - (float)j0, fz = cz - (float)k0;
        int ex, ey, ez;
        face_extent(comp, a.d, ex, ey, ez);
        const float vc = a.vel[3 * p + comp];
        float vbar = 0.0f;
        for (int c = 0; c < 8; ++c) {
            const int di = c >> 2, dj = (c >> 1) & 1, dk = c & 1;
            const int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
            if (ii < 0 || ii >= ex || jj < 0 || jj >= ey || kk < 0 || kk >= ez) continue;
            const size_t idx = face_index(comp, ii, jj, kk, a.d);
            const float wx = di ? fx : 1.0f - fx, wy = dj ? fy : 1.0f - fy, wz = dk ? fz : 1.0f - fz;
            const float qb = a.qb[comp][idx], mb = a.mb[comp][idx];
            vbar += wx * wy * wz * qb;
            const float t = qb * vc + mb;
            gx += (di ? 1.0f : -1.0f) * wy * wz * t;
            gy += wx * (dj ? 1.0f : -1.0f) * wz * t;
            gz += wx * wy * (dk ? 1.0f : -1.0f) * t;
        }
        a.avel[3 * id + comp] += vbar;
    }
    a.apos[3 * id] += gx * inv;
    a.apos[3 * id + 1] += gy * inv;
    a.apos[3 * id + 2] += gz * inv;
}

AdjDims dims_of(const gf_adjoint_t* a) {
    AdjDims d;
    d.nx = a->nx; d.ny = a->ny; d.nz = a->nz; d.dx = a->dx;
    return d;
}

int adj_check(const gf_adjoint_t* a, const char* who) {
    if (!a) return gf::set_error(GF_EINVAL, "%s: null descriptor", who);
    if (a->nx < 1 || a->ny < 1 || a->nz < 1 || !(a->dx > 0.0f))
        return gf::set_error(GF_EINVAL, "%s: grid %dx%dx%d dx %g", who, a->nx, a->ny, a->nz,
                             (double)a->dx);
    if (a->n < 0) return gf::set_error(GF_EINVAL, "%s: n %d", who, a->n);
    if (a->n > 0 && (!a->pos || !a->vel || !a->pid || !a->apos || !a->avel))
        return gf::set_error(GF_EINVAL, "%s: particle buffers not set", who);
    return 0;
}

size_t face_count(const gf_adjoint_t* a, int comp) {
    return (size_t)(a->nx + (comp == 0)) * (a->ny + (comp == 1)) * (a->nz + (comp == 2));
}

}  // namespace

extern "C" int gf_adj_clear(const gf_adjoint_t* a, void* stream) {
    int rc = adj_check(a, "gf_adj_clear");
    if (rc) return rc;
    float* bufs[9] = {a->mu, a->mv, a->mw, a->au, a->av, a->aw, a->aus, a->avs, a->aws};
    for (int b = 0; b < 9; ++b) GF_REQUIRE(bufs[b], "gf_adj_clear: face buffers not set");
    hipStream_t st = (hipStream_t)stream;
    for (int b = 0; b < 9; ++b) {
        const size_t n = face_count(a, b % 3);
        hipLaunchKernelGGL(k_adj_fill, dim3((unsigned)((n + GF_WG - 1) / GF_WG)), dim3(GF_WG), 0, st,
                           bufs[b], n);
    }
    GF_LAUNCH_CHECK("k_adj_fill");
    return 0;
}

extern "C" int gf_adj_weights(const gf_adjoint_t* a, void* stream) {
    int rc = adj_check(a, "gf_adj_weights");
    if (rc) return rc;
    GF_REQUIRE(a->mu && a->mv && a->mw, "gf_adj_weights: weight buffers not set");
    if (a->n == 0) return 0;
    hipLaunchKernelGGL(k_adj_weights, dim3((a->n + GF_WG - 1) / GF_WG), dim3(GF_WG), 0,
                       (hipStream_t)stream, a->pos, a->n, dims_of(a), a->mu, a->mv, a->mw);
    GF_LAUNCH_CHECK("k_adj_weights");
    return 0;
}

extern "C" int gf_adj_g2p(const gf_adjoint_t* a, float flip_blend, float dt, void* stream) {
    int rc = adj_check(a, "gf_adj_g2p");
    if (rc) return rc;
    GF_REQUIRE(a->u && a->v && a->w && a->us && a->vs && a->ws && a->au && a->av && a->aw &&
                   a->aus && a->avs && a->aws,
               "gf_adj_g2p: face buffers not set");
    if (a->n == 0) return 0;
    G2PAdj g;
    g.pos = a->pos; g.vel = a->vel; g.pid = a->pid;
    g.f[0] = a->u; g.f[1] = a->v; g.f[2] = a->w;
    g.fs[0] = a->us; g.fs[1] = a->vs; g.fs[2] = a->ws;
    g.af[0] = a->au; g.af[1] = a->av; g.af[2] = a->aw;
    g.afs[0] = a->aus; g.afs[1] = a->avs; g.afs[2] = a->aws;
    g.apos = a->apos; g.avel = a->avel; g.n = a->n; g.d = dims_of(a);
    g.fb = flip_blend; g.dt = dt;
    g.dom[0] = a->dom[0]; g.dom[1] = a->dom[1]; g.dom[2] = a->dom[2];
    hipLaunchKernelGGL(k_adj_g2p, dim3((a->n + GF_WG - 1) / GF_WG), dim3(GF_WG), 0,
                       (hipStream_t)stream, g);
    GF_LAUNCH_CHECK("k_adj_g2p");
    return 0;
}

extern "C" int gf_adj_pressure_rhs(const gf_adjoint_t* a, float dt, float rho, float* pbar,
                                   void* stream) {
    int rc = adj_check(a, "gf_adj_pressure_rhs");
    if (rc) return rc;
    GF_REQUIRE(a->marker && a->tile_map && a->au && a->av && a->aw && pbar,
               "gf_adj_pressure_rhs: buffers not set");
    GF_REQUIRE(rho != 0.0f, "gf_adj_pressure_rhs: rho must be nonzero");
    const size_t C = (size_t)a->nx * a->ny * a->nz;
    const float sg = dt / (rho * a->dx);
    hipLaunchKernelGGL(k_adj_pressure_rhs, dim3((unsigned)((C + GF_WG - 1) / GF_WG)), dim3(GF_WG), 0,
                       (hipStream_t)stream, a->marker, a->tile_map, a->au, a->av, a->aw,
                       dims_of(a), sg, pbar);
    GF_LAUNCH_CHECK("k_adj_pressure_rhs");
    return 0;
}

extern "C" int gf_adj_faces(const gf_adjoint_t* a, const float* bbar, float dt, float rho,
                            void* stream) {
    int rc = adj_check(a, "gf_adj_faces");
    if (rc) return rc;
    GF_REQUIRE(a->marker && a->tile_map && bbar && a->mu && a->mv && a->mw && a->us && a->vs &&
                   a->ws && a->au && a->av && a->aw && a->aus && a->avs && a->aws,
               "gf_adj_faces: buffers not set");
    GF_REQUIRE(dt != 0.0f, "gf_adj_faces: dt must be nonzero");
    const float sd = rho * a->dx / dt;
    const AdjDims d = dims_of(a);
    hipStream_t st = (hipStream_t)stream;
    const size_t n0 = face_count(a, 0), n1 = face_count(a, 1), n2 = face_count(a, 2);
    hipLaunchKernelGGL(k_adj_faces<0>, dim3((unsigned)((n0 + GF_WG - 1) / GF_WG)), dim3(GF_WG), 0, st,
                       a->marker, a->tile_map, bbar, a->mu, a->us, a->au, a->aus, d, sd);
    hipLaunchKernelGGL(k_adj_faces<1>, dim3((unsigned)((n1 + GF_WG - 1) / GF_WG)), dim3(GF_WG), 0, st,
                       a->marker, a->tile_map, bbar, a->mv, a->vs, a->av, a->avs, d, sd);
    hipLaunchKernelGGL(k_adj_faces<2>, dim3((unsigned)((n2 + GF_WG - 1) / GF_WG)), dim3(GF_WG), 0, st,
                       a->marker, a->tile_map, bbar, a->mw, a->ws, a->aw, a->aws, d, sd);
    GF_LAUNCH_CHECK("k_adj_faces");
    return 0;
}

extern "C" int gf_adj_p2g(const gf_adjoint_t* a, void* stream) {
    int rc = adj_check(a, "gf_adj_p2g");
    if (rc) return rc;
    GF_REQUIRE(a->au && a->av && a->aw && a->aus && a->avs && a->aws,
               "gf_adj_p2g: face buffers not set");
    if (a->n == 0) return 0;
    P2GAdj g;
    g.pos = a->pos; g.vel = a->vel; g.pid = a->pid;
    g.qb[0] = a->au; g.qb[1] = a->av; g.qb[2] = a->aw;
    g.mb[0] = a->aus; g.mb[1] = a->avs; g.mb[2] = a->aws;
    g.apos = a->apos; g.avel = a->avel; g.n = a->n; g.d = dims_of(a);
    hipLaunchKernelGGL(k_adj_p2g, dim3((a->n + GF_WG - 1) / GF_WG), dim3(GF_WG), 0,
                       (hipStream_t)stream, g);
    GF_LAUNCH_CHECK("k_adj_p2g");
    return 0;
}
