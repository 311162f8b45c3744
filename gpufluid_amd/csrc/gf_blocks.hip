// Dense-grid replacements of the reference's individual S2 kernels, with the
// reference argument meaning (sub-dense offsets fixed at 0).  They serve
// callers that launch one block at a time (the S2 block interface); the
// substep itself runs the fused tile kernels of gf_transfer.hip and
// gf_pressure.hip.
#include "gf_common.h"

namespace {

__device__ __forceinline__ bool idx3(int I, int J, int K, int& i, int& j, int& k) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)I * J * K) return false;
    k = (int)(t % K);
    j = (int)((t / K) % J);
    i = (int)(t / ((long long)K * J));
    return true;
}

inline dim3 grid_for(long long n) { return dim3((unsigned)((n + GF_WG - 1) / GF_WG)); }

// Cells a pressure block kernel visits: every cell (coords NULL), or the
// per-tile forms' cells (solver3d.py:2020-2035, 1120-1160): thread t covers
// cell (di, dj, dk) = t mod bs^3 (dk fastest) of listed block t / bs^3, for
// the first n_active[0] blocks of coords (block coordinates, 3 ints each),
// skipping cells outside the grid.
struct Dom {
    const int32_t* coords;
    const int32_t* n_active;
    int bs;
};

__device__ __forceinline__ bool dom_cell(const Dom& d, int nx, int ny, int nz, int& i, int& j,
                                         int& k) {
    if (!d.coords) return idx3(nx, ny, nz, i, j, k);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long cpb = (long long)d.bs * d.bs * d.bs;
    const long long blk = t / cpb;
    if (blk >= (long long)d.n_active[0]) return false;
    const int rem = (int)(t - blk * cpb);
    i = d.coords[3 * blk + 0] * d.bs + rem / (d.bs * d.bs);
    j = d.coords[3 * blk + 1] * d.bs + (rem / d.bs) % d.bs;
    k = d.coords[3 * blk + 2] * d.bs + rem % d.bs;
    return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz;
}

// Block-sparse skip of k3_jacobi_pressure_blocksparse (solver3d.py:975-1041):
// cells of blocks with active[bi, bj, bk] == 0 get p_out = 0.
struct BlkMask {
    const int32_t* active;  // (nbx, nby, nbz) or NULL
    int bs, nby, nbz;
};

// solver3d.py:41-65
__global__ void k_clear_grid(float* u, float* v, float* w, float* uw, float* vw, float* ww,
                             int32_t* marker, int nx, int ny, int nz) {
    int i, j, k;
    if (!idx3(nx + 1, ny + 1, nz + 1, i, j, k)) return;
    if (i < nx + 1 && j < ny && k < nz) {
        size_t f = ((size_t)i * ny + j) * nz + k;
        u[f] = 0.0f; uw[f] = 0.0f;
    }
    if (i < nx && j < ny + 1 && k < nz) {
        size_t f = ((size_t)i * (ny + 1) + j) * nz + k;
        v[f] = 0.0f; vw[f] = 0.0f;
    }
    if (i < nx && j < ny && k < nz + 1) {
        size_t f = ((size_t)i * ny + j) * (nz + 1) + k;
        w[f] = 0.0f; ww[f] = 0.0f;
    }
    if (i < nx && j < ny && k < nz) {
        size_t c = ((size_t)i * ny + j) * nz + k;
        if (marker[c] != 2) marker[c] = 0;
    }
}

// solver3d.py:96-101
__global__ void k_mark_fluid(const float* pos, int n, int32_t* marker, float dx, int nx, int ny,
                             int nz) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int ci = gf::bin_axis(pos[3 * p + 0], dx, nx);
    int cj = gf::bin_axis(pos[3 * p + 1], dx, ny);
    int ck = gf::bin_axis(pos[3 * p + 2], dx, nz);
    size_t c = ((size_t)ci * ny + cj) * nz + ck;
    if (marker[c] != 2) marker[c] = 1;
}

// solver3d.py:121-142
__global__ void k_normalize(float* u, float* v, float* w, const float* uw, const float* vw,
                            const float* ww, float* us, float* vs, float* ws, int nx, int ny,
                            int nz) {
    int i, j, k;
    if (!idx3(nx + 1, ny + 1, nz + 1, i, j, k)) return;
    if (i < nx + 1 && j < ny && k < nz) {
        size_t f = ((size_t)i * ny + j) * nz + k;
        float val = uw[f] > 1.0e-8f ? u[f] / uw[f] : 0.0f;
        u[f] = val; us[f] = val;
    }
    if (i < nx && j < ny + 1 && k < nz) {
        size_t f = ((size_t)i * (ny + 1) + j) * nz + k;
        float val = vw[f] > 1.0e-8f ? v[f] / vw[f] : 0.0f;
        v[f] = val; vs[f] = val;
    }
    if (i < nx && j < ny && k < nz + 1) {
        size_t f = ((size_t)i * ny + j) * (nz + 1) + k;
        float val = ww[f] > 1.0e-8f ? w[f] / ww[f] : 0.0f;
        w[f] = val; ws[f] = val;
    }
}

// solver3d.py:149-156
__global__ void k_add_gravity(float* v, long long n, float g, float dt) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = v[i] + g * dt;
}

// solver3d.py:758-799
__global__ void k_enforce_solid_bc(float* u, float* v, float* w, const int32_t* marker,
                                   const float* su, const float* sv, const float* sw, int nx,
                                   int ny, int nz) {
    // 64 lanes along k x 4 along j per workgroup, one i per grid z
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k > nz || j > ny || i > nx) return;
    if (i <= nx && j < ny && k < nz) {
        bool left = (i == 0) || (i > 0 && marker[((size_t)(i - 1) * ny + j) * nz + k] == 2);
        bool right = (i == nx) || (i < nx && marker[((size_t)i * ny + j) * nz + k] == 2);
        if (left || right) {
            size_t f = ((size_t)i * ny + j) * nz + k;
            u[f] = su ? su[f] : 0.0f;
        }
    }
    if (i < nx && j <= ny && k < nz) {
        bool bot = (j == 0) || (j > 0 && marker[((size_t)i * ny + j - 1) * nz + k] == 2);
        bool top = (j == ny) || (j < ny && marker[((size_t)i * ny + j) * nz + k] == 2);
        if (bot || top) {
            size_t f = ((size_t)i * (ny + 1) + j) * nz + k;
            v[f] = sv ? sv[f] : 0.0f;
        }
    }
    if (i < nx && j < ny && k <= nz) {
        bool back = (k == 0) || (k > 0 && marker[((size_t)i * ny + j) * nz + k - 1] == 2);
        bool front = (k == nz) || (k < nz && marker[((size_t)i * ny + j) * nz + k] == 2);
        if (back || front) {
            size_t f = ((size_t)i * ny + j) * (nz + 1) + k;
            w[f] = sw ? sw[f] : 0.0f;
        }
    }
}

// solver3d.py:846-871
__global__ void k_divergence(const float* u, const float* v, const float* w, float* div,
                             const int32_t* marker, int nx, int ny, int nz, float dx, float dt,
                             float rho) {
    int i, j, k;
    if (!idx3(nx, ny, nz, i, j, k)) return;
    size_t c = ((size_t)i * ny + j) * nz + k;
    if (marker[c] != 1) { div[c] = 0.0f; return; }
    float u0 = u[((size_t)i * ny + j) * nz + k], u1 = u[((size_t)(i + 1) * ny + j) * nz + k];
    float v0 = v[((size_t)i * (ny + 1) + j) * nz + k], v1 = v[((size_t)i * (ny + 1) + j + 1) * nz + k];
    float w0 = w[((size_t)i * ny + j) * (nz + 1) + k], w1 = w[((size_t)i * ny + j) * (nz + 1) + k + 1];
    float d = u1 - u0 + v1 - v0 + w1 - w0;
    div[c] = d * (rho * dx / dt);
}

// Neighbour pass shared by the Jacobi sweep and A.x (solver3d.py:878-938, 1743-1797).
__device__ __forceinline__ void stencil(const float* x, const int32_t* marker, int i, int j,
                                        int k, int nx, int ny, int nz, float& sum_nb,
                                        float& diag) {
    const size_t c = ((size_t)i * ny + j) * nz + k;
    const size_t sx = (size_t)ny * nz, sy = nz;
    sum_nb = 0.0f; diag = 0.0f;
    int m;
    if (i > 0) { m = marker[c - sx]; if (m != 2) { diag += 1.0f; if (m == 1) sum_nb += x[c - sx]; } }
    if (i < nx - 1) { m = marker[c + sx]; if (m != 2) { diag += 1.0f; if (m == 1) sum_nb += x[c + sx]; } }
    if (j > 0) { m = marker[c - sy]; if (m != 2) { diag += 1.0f; if (m == 1) sum_nb += x[c - sy]; } }
    if (j < ny - 1) { m = marker[c + sy]; if (m != 2) { diag += 1.0f; if (m == 1) sum_nb += x[c + sy]; } }
    if (k > 0) { m = marker[c - 1]; if (m != 2) { diag += 1.0f; if (m == 1) sum_nb += x[c - 1]; } }
    if (k < nz - 1) { m = marker[c + 1]; if (m != 2) { diag += 1.0f; if (m == 1) sum_nb += x[c + 1]; } }
}

__global__ void k_jacobi_dense(const float* p_in, float* p_out, const float* div,
                               const int32_t* marker, int nx, int ny, int nz, Dom dom,
                               BlkMask bm) {
    int i, j, k;
    if (!dom_cell(dom, nx, ny, nz, i, j, k)) return;
    size_t c = ((size_t)i * ny + j) * nz + k;
    if (bm.active &&
        bm.active[((size_t)(i / bm.bs) * bm.nby + j / bm.bs) * bm.nbz + k / bm.bs] == 0) {
        p_out[c] = 0.0f;
        return;
    }
    if (marker[c] != 1) { p_out[c] = 0.0f; return; }
    float s, d;
    stencil(p_in, marker, i, j, k, nx, ny, nz, s, d);
    if (d < 0.5f) { p_out[c] = 0.0f; return; }
    p_out[c] = (s - div[c]) / d;
}

// One red-black Gauss-Seidel half sweep in place (solver3d.py:1204-1267):
// cells with (i+j+k) % 2 == color.
__global__ void k_gsrb_dense(float* p, const float* div, const int32_t* marker, int nx, int ny,
                             int nz, int color, Dom dom) {
    int i, j, k;
    if (!dom_cell(dom, nx, ny, nz, i, j, k)) return;
    if (((i + j + k) & 1) != color) return;
    size_t c = ((size_t)i * ny + j) * nz + k;
    if (marker[c] != 1) { p[c] = 0.0f; return; }
    float s, d;
    stencil(p, marker, i, j, k, nx, ny, nz, s, d);
    p[c] = d < 0.5f ? 0.0f : (s - div[c]) / d;
}

// Products and sums rounded one by one (no fused multiply-add), as the FP32
// oracle computes them.
__global__ void k_apply_A_dense(const float* x, float* y, const int32_t* marker,
                                const int32_t* done, int nx, int ny, int nz, Dom dom) {
#pragma clang fp contract(off)
    if (done && done[0] != 0) return;
    int i, j, k;
    if (!dom_cell(dom, nx, ny, nz, i, j, k)) return;
    size_t c = ((size_t)i * ny + j) * nz + k;
    if (marker[c] != 1) { y[c] = 0.0f; return; }
    float s, d;
    stencil(x, marker, i, j, k, nx, ny, nz, s, d);
    y[c] = d * x[c] - s;
}

__global__ void k_compute_diag_dense(const int32_t* marker, float* diag, int nx, int ny, int nz) {
    int i, j, k;
    if (!idx3(nx, ny, nz, i, j, k)) return;
    size_t c = ((size_t)i * ny + j) * nz + k;
    if (marker[c] != 1) { diag[c] = 1.0f; return; }
    const size_t sx = (size_t)ny * nz, sy = nz;
    float d = 0.0f;
    if (i > 0 && marker[c - sx] != 2) d += 1.0f;
    if (i < nx - 1 && marker[c + sx] != 2) d += 1.0f;
    if (j > 0 && marker[c - sy] != 2) d += 1.0f;
    if (j < ny - 1 && marker[c + sy] != 2) d += 1.0f;
    if (k > 0 && marker[c - 1] != 2) d += 1.0f;
    if (k < nz - 1 && marker[c + 1] != 2) d += 1.0f;
    diag[c] = fmaxf(d, 1.0f);
}

__global__ void k_apply_invM_dense(const float* r, const float* diag, float* z,
                                   const int32_t* marker, const int32_t* done, int nx, int ny,
                                   int nz, Dom dom) {
    if (done && done[0] != 0) return;
    int i, j, k;
    if (!dom_cell(dom, nx, ny, nz, i, j, k)) return;
    const size_t c = ((size_t)i * ny + j) * nz + k;
    z[c] = marker[c] != 1 ? 0.0f : r[c] / diag[c];
}

__global__ __launch_bounds__(GF_WG) void k_dot_dense(const float* a, const float* b, float* out,
                                                     const int32_t* marker, const int32_t* done,
                                                     int nx, int ny, int nz, Dom dom) {
    __shared__ double s_w[GF_WG / GF_WAVE];
    if (done && done[0] != 0) return;
    int i, j, k;
    double v = 0.0;
    if (dom_cell(dom, nx, ny, nz, i, j, k)) {
        const size_t c = ((size_t)i * ny + j) * nz + k;
        if (marker[c] == 1) v = (double)(a[c] * b[c]);
    }
    v = gf::wave_sum(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < GF_WG / GF_WAVE; ++i) t += s_w[i];
        if (t != 0.0) atomicAdd(out, (float)t);
    }
}

__global__ void k_axpy_devscalar_dense(const float* a_dev, float sign, const float* x,
                                       const float* y_in, float* y_out, const int32_t* marker,
                                       const int32_t* done, int nx, int ny, int nz, Dom dom) {
#pragma clang fp contract(off)
    if (done && done[0] != 0) return;
    int i, j, k;
    if (!dom_cell(dom, nx, ny, nz, i, j, k)) return;
    const size_t c = ((size_t)i * ny + j) * nz + k;
    if (marker[c] != 1) { y_out[c] = 0.0f; return; }
    float a = a_dev[0] * sign;
    y_out[c] = y_in[c] + a * x[c];
}

// k_mark_active_blocks (solver3d.py:952-969): a fluid cell sets its block's
// flag to 1 (every writer stores the same value); other flags are untouched.
__global__ void k_mark_blocks(const int32_t* marker, int32_t* active, int nx, int ny, int nz,
                              int bs, int nby, int nbz) {
    int i, j, k;
    if (!idx3(nx, ny, nz, i, j, k)) return;
    if (marker[((size_t)i * ny + j) * nz + k] != 1) return;
    active[((size_t)(i / bs) * nby + j / bs) * nbz + k / bs] = 1;
}

__global__ void k_scalar_op(int op, float* a, float* b, float* c, float tol, int32_t* done) {
    switch (op) {
        case 0:  // k3_zero_scalar(s, done)
            if (done && done[0] != 0) return;
            a[0] = 0.0f;
            break;
        case 1: {  // k3_div_scalar(numer, denom, out, done)
            if (done && done[0] != 0) return;
            float d = b[0];
            c[0] = fabsf(d) < 1.0e-30f ? 0.0f : a[0] / d;
            break;
        }
        case 2:  // k3_copy_scalar(src, dst, done)
            if (done && done[0] != 0) return;
            c[0] = a[0];
            break;
        case 3:  // k3_set_tol_sq(r0_norm2, tol, tol_sq)
            c[0] = tol * tol * a[0];
            break;
        case 4:  // k3_check_converged(r_norm2, tol_sq, done)
            if (a[0] < b[0]) done[0] = 1;
            break;
        case 5:  // k3_zero_int_scalar(s)
            done[0] = 0;
            break;
    }
}

// solver3d.py:1364-1407
__global__ void k_subtract_grad(float* u, float* v, float* w, const float* p,
                                const int32_t* marker, int nx, int ny, int nz, float dx,
                                float dt, float rho) {
    int i, j, k;
    if (!idx3(nx + 1, ny + 1, nz + 1, i, j, k)) return;
    const float scale = dt / (rho * dx);
    auto M = [&](int a, int b2, int c) { return marker[((size_t)a * ny + b2) * nz + c]; };
    auto P = [&](int a, int b2, int c) { return p[((size_t)a * ny + b2) * nz + c]; };
    if (i >= 1 && i <= nx - 1 && j < ny && k < nz) {
        int ml = M(i - 1, j, k), mr = M(i, j, k);
        if ((ml == 1 || mr == 1) && ml != 2 && mr != 2) {
            float pl = ml == 1 ? P(i - 1, j, k) : 0.0f;
            float pr = mr == 1 ? P(i, j, k) : 0.0f;
            size_t f = ((size_t)i * ny + j) * nz + k;
            u[f] = u[f] - scale * (pr - pl);
        }
    }
    if (j >= 1 && j <= ny - 1 && i < nx && k < nz) {
        int mb = M(i, j - 1, k), mt = M(i, j, k);
        if ((mb == 1 || mt == 1) && mb != 2 && mt != 2) {
            float pb = mb == 1 ? P(i, j - 1, k) : 0.0f;
            float pt = mt == 1 ? P(i, j, k) : 0.0f;
            size_t f = ((size_t)i * (ny + 1) + j) * nz + k;
            v[f] = v[f] - scale * (pt - pb);
        }
    }
    if (k >= 1 && k <= nz - 1 && i < nx && j < ny) {
        int mba = M(i, j, k - 1), mfr = M(i, j, k);
        if ((mba == 1 || mfr == 1) && mba != 2 && mfr != 2) {
            float pba = mba == 1 ? P(i, j, k - 1) : 0.0f;
            float pfr = mfr == 1 ? P(i, j, k) : 0.0f;
            size_t f = ((size_t)i * ny + j) * (nz + 1) + k;
            w[f] = w[f] - scale * (pfr - pba);
        }
    }
}

// gridmath.py:30-65 on a dense face array of shape (nmx, nmy, nmz).
__device__ float sample3(const float* f, float fx, float fy, float fz, int nmx, int nmy, int nmz) {
#pragma clang fp contract(off)
    fx = fminf(fmaxf(fx, 0.0f), (float)nmx);
    fy = fminf(fmaxf(fy, 0.0f), (float)nmy);
    fz = fminf(fmaxf(fz, 0.0f), (float)nmz);
    int i0 = (int)floorf(fx), j0 = (int)floorf(fy), k0 = (int)floorf(fz);
    if (i0 >= nmx) i0 = nmx - 1;
    if (j0 >= nmy) j0 = nmy - 1;
    if (k0 >= nmz) k0 = nmz - 1;
    if (i0 < 0) i0 = 0;
    if (j0 < 0) j0 = 0;
    if (k0 < 0) k0 = 0;
    int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
    if (i1 >= nmx) i1 = nmx - 1;
    if (j1 >= nmy) j1 = nmy - 1;
    if (k1 >= nmz) k1 = nmz - 1;
    const float sx = fminf(fmaxf(fx - (float)i0, 0.0f), 1.0f);
    const float sy = fminf(fmaxf(fy - (float)j0, 0.0f), 1.0f);
    const float sz = fminf(fmaxf(fz - (float)k0, 0.0f), 1.0f);
    auto F = [&](int a, int b, int c) { return f[((size_t)a * nmy + b) * nmz + c]; };
    const float c000 = F(i0, j0, k0), c100 = F(i1, j0, k0);
    const float c010 = F(i0, j1, k0), c110 = F(i1, j1, k0);
    const float c001 = F(i0, j0, k1), c101 = F(i1, j0, k1);
    const float c011 = F(i0, j1, k1), c111 = F(i1, j1, k1);
    const float c00 = c000 * (1.0f - sx) + c100 * sx;
    const float c10 = c010 * (1.0f - sx) + c110 * sx;
    const float c01 = c001 * (1.0f - sx) + c101 * sx;
    const float c11 = c011 * (1.0f - sx) + c111 * sx;
    const float c0 = c00 * (1.0f - sy) + c10 * sy;
    const float c1 = c01 * (1.0f - sy) + c11 * sy;
    return c0 * (1.0f - sz) + c1 * sz;
}

// solver3d.py:1414-1463
__global__ void k_g2p_advect_dense(float* pos, float* vel, const float* u, const float* v,
                                   const float* w, const float* us, const float* vs,
                                   const float* ws, int n, float dx, float dt, int nx, int ny,
                                   int nz, float fb, float dmx, float dmy, float dmz) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    const float lx = px / dx, ly = py / dx, lz = pz / dx;
    const float pic_u = sample3(u, lx, ly - 0.5f, lz - 0.5f, nx + 1, ny, nz);
    const float old_u = sample3(us, lx, ly - 0.5f, lz - 0.5f, nx + 1, ny, nz);
    const float pic_v = sample3(v, lx - 0.5f, ly, lz - 0.5f, nx, ny + 1, nz);
    const float old_v = sample3(vs, lx - 0.5f, ly, lz - 0.5f, nx, ny + 1, nz);
    const float pic_w = sample3(w, lx - 0.5f, ly - 0.5f, lz, nx, ny, nz + 1);
    const float old_w = sample3(ws, lx - 0.5f, ly - 0.5f, lz, nx, ny, nz + 1);
    const float fu = vel[3 * p] + (pic_u - old_u);
    const float fv = vel[3 * p + 1] + (pic_v - old_v);
    const float fw = vel[3 * p + 2] + (pic_w - old_w);
    const float nu = fb * fu + (1.0f - fb) * pic_u;
    const float nv = fb * fv + (1.0f - fb) * pic_v;
    const float nw = fb * fw + (1.0f - fb) * pic_w;
    vel[3 * p] = nu; vel[3 * p + 1] = nv; vel[3 * p + 2] = nw;
    const float eps = dx * 1.001f;
    pos[3 * p] = fminf(fmaxf(px + nu * dt, eps), dmx - eps);
    pos[3 * p + 1] = fminf(fmaxf(py + nv * dt, eps), dmy - eps);
    pos[3 * p + 2] = fminf(fmaxf(pz + nw * dt, eps), dmz - eps);
}

// solver3d.py:1591-1732
template <int COMP>
__device__ void apic_dense(const float* f, float px, float py, float pz, float dx,
                           float inv_dx2_3, int nx, int ny, int nz, float& vout, float* crow) {
    float fx = px / dx, fy = py / dx, fz = pz / dx;
    if (COMP != 0) fx = fx - 0.5f;
    if (COMP != 1) fy = fy - 0.5f;
    if (COMP != 2) fz = fz - 0.5f;
    const int i0 = (int)floorf(fx), j0 = (int)floorf(fy), k0 = (int)floorf(fz);
    const float sx = fx - (float)i0, sy = fy - (float)j0, sz = fz - (float)k0;
    const int imax = COMP == 0 ? nx : nx - 1, jmax = COMP == 1 ? ny : ny - 1,
              kmax = COMP == 2 ? nz : nz - 1;
    const int nfj = ny + (COMP == 1), nfk = nz + (COMP == 2);
    float acc = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0) ii = 0;
                if (jj < 0) jj = 0;
                if (kk < 0) kk = 0;
                if (ii > imax) ii = imax;
                if (jj > jmax) jj = jmax;
                if (kk > kmax) kk = kmax;
                const float wt = (di ? sx : 1.0f - sx) * (dj ? sy : 1.0f - sy) * (dk ? sz : 1.0f - sz);
                const float val = f[((size_t)ii * nfj + jj) * nfk + kk];
                acc += wt * val;
                const float ox = (COMP == 0 ? (float)ii : (float)ii + 0.5f) * dx - px;
                const float oy = (COMP == 1 ? (float)jj : (float)jj + 0.5f) * dx - py;
                const float oz = (COMP == 2 ? (float)kk : (float)kk + 0.5f) * dx - pz;
                const float fu = wt * val * inv_dx2_3;
                c0 += fu * ox; c1 += fu * oy; c2 += fu * oz;
            }
    vout = acc; crow[0] = c0; crow[1] = c1; crow[2] = c2;
}

__global__ void k_g2p_apic_dense(float* pos, float* vel, float* C, const float* u, const float* v,
                                 const float* w, int n, float dx, float dt, int nx, int ny,
                                 int nz, float dmx, float dmy, float dmz) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
    const float inv = 3.0f / (dx * dx);
    float nu, nv, nw, cm[9];
    apic_dense<0>(u, px, py, pz, dx, inv, nx, ny, nz, nu, cm);
    apic_dense<1>(v, px, py, pz, dx, inv, nx, ny, nz, nv, cm + 3);
    apic_dense<2>(w, px, py, pz, dx, inv, nx, ny, nz, nw, cm + 6);
    vel[3 * p] = nu; vel[3 * p + 1] = nv; vel[3 * p + 2] = nw;
    for (int c = 0; c < 9; ++c) C[9 * p + c] = cm[c];
    const float eps = dx * 1.001f;
    pos[3 * p] = fminf(fmaxf(px + nu * dt, eps), dmx - eps);
    pos[3 * p + 1] = fminf(fmaxf(py + nv * dt, eps), dmy - eps);
    pos[3 * p + 2] = fminf(fmaxf(pz + nw * dt, eps), dmz - eps);
}

// solver3d.py:2241-2252 (non-negative floats order like their bit patterns)
__global__ __launch_bounds__(GF_WG) void k_vmax_sq(const float* vel, int n, float* out) {
    __shared__ float s_w[GF_WG / GF_WAVE];
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (p < n) {
        float a = vel[3 * p], b = vel[3 * p + 1], c = vel[3 * p + 2];
        s = a * a + b * b + c * c;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s = fmaxf(s, __shfl_xor(s, off, 64));
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int i = 0; i < GF_WG / GF_WAVE; ++i) m = fmaxf(m, s_w[i]);
        atomicMax((unsigned int*)out, __float_as_uint(m));
    }
}

// k3_jacobi_visc (solver3d.py:178-197): one Jacobi sweep of
// (I - r Lap) u_new = u_old on one MAC component of shape (sx, sy, sz).
// Neighbours in range only, summed in the order -x, +x, -y, +y, -z, +z;
// u_out = (u_old + r*sum) / (1 + r*count) with every operation rounded on
// its own.  Threads: 64 along k x 4 along j per workgroup, one i per grid z.
__global__ __launch_bounds__(GF_WG) void k_jacobi_visc(const float* __restrict__ u_old,
                                                       const float* __restrict__ u_in,
                                                       float* __restrict__ u_out, int sx,
                                                       int sy, int sz, float r) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= sz || j >= sy) return;
    const size_t di = (size_t)sy * sz, dj = (size_t)sz;
    const size_t c = (size_t)i * di + (size_t)j * dj + k;
    float sum = 0.0f, cnt = 0.0f;
    if (i > 0) { sum += u_in[c - di]; cnt += 1.0f; }
    if (i < sx - 1) { sum += u_in[c + di]; cnt += 1.0f; }
    if (j > 0) { sum += u_in[c - dj]; cnt += 1.0f; }
    if (j < sy - 1) { sum += u_in[c + dj]; cnt += 1.0f; }
    if (k > 0) { sum += u_in[c - 1]; cnt += 1.0f; }
    if (k < sz - 1) { sum += u_in[c + 1]; cnt += 1.0f; }
    const float num = u_old[c] + r * sum;
    const float den = 1.0f + r * cnt;
    u_out[c] = num / den;
}

// The same sweep on all three MAC components in one launch, for the substep's
// viscosity stage.  A thread owns one (j, k) column of one component and
// marches along i through a chunk of VISC_CHUNK planes, VISC_BATCH planes at
// a time: the batch's loads are all issued before its stores, and the i-1,
// i, i+1 values of u_in stay in registers, so u_in is read about once.
constexpr int VISC_CHUNK = 32;
constexpr int VISC_BATCH = 4;

// The sweep runs over a box of cells [o, o + b) of the nx*ny*nz grid (the
// whole grid, or the sub-dense box of solver3d.py:3821-3906, where the
// reference's k3_jacobi_visc runs on the box-shaped arrays): a component's
// faces of the box, with the box as the array bounds of the stencil.
struct ViscArgs {
    const float* old0; const float* old1; const float* old2;
    const float* in0; const float* in1; const float* in2;
    float* out0; float* out1; float* out2;
    int nx, ny, nz;   // grid cells (strides)
    int ox, oy, oz;   // box origin (cells)
    int bx, by, bz;   // box extent (cells)
    int nchunk;
    float r;
};

// One (j, k) column of a face array: sx, sy, sz are the stencil bounds (the
// box's face extents), di / dj the strides of the full array, the pointers
// already at the box origin.
__device__ __forceinline__ void visc_column(const float* __restrict__ u_old,
                                            const float* __restrict__ u_in,
                                            float* __restrict__ u_out, int sx, int sy, int sz,
                                            size_t di, size_t dj, int i0, int i1, int j, int k,
                                            float r) {
#pragma clang fp contract(off)
    const size_t base = (size_t)j * dj + k;
    // col[m] = u_in at plane i + m - 1 of the current batch (m = 0..VISC_BATCH+1)
    float col[VISC_BATCH + 2];
    col[0] = i0 > 0 ? u_in[(size_t)(i0 - 1) * di + base] : 0.0f;
    col[1] = u_in[(size_t)i0 * di + base];
    for (int i = i0; i < i1; i += VISC_BATCH) {
        float ym[VISC_BATCH], yp[VISC_BATCH], zm[VISC_BATCH], zp[VISC_BATCH], uo[VISC_BATCH];
#pragma unroll
        for (int b = 0; b < VISC_BATCH; ++b) {
            const int ii = i + b;
            const size_t q = (size_t)ii * di + base;
            col[b + 2] = ii + 1 < sx ? u_in[q + di] : 0.0f;
            const bool live = ii < i1;
            ym[b] = live && j > 0 ? u_in[q - dj] : 0.0f;
            yp[b] = live && j < sy - 1 ? u_in[q + dj] : 0.0f;
            zm[b] = live && k > 0 ? u_in[q - 1] : 0.0f;
            zp[b] = live && k < sz - 1 ? u_in[q + 1] : 0.0f;
            uo[b] = live ? u_old[q] : 0.0f;
        }
#pragma unroll
        for (int b = 0; b < VISC_BATCH; ++b) {
            const int ii = i + b;
            if (ii >= i1) break;
            float sum = 0.0f, cnt = 0.0f;
            if (ii > 0) { sum += col[b]; cnt += 1.0f; }
            if (ii < sx - 1) { sum += col[b + 2]; cnt += 1.0f; }
            if (j > 0) { sum += ym[b]; cnt += 1.0f; }
            if (j < sy - 1) { sum += yp[b]; cnt += 1.0f; }
            if (k > 0) { sum += zm[b]; cnt += 1.0f; }
            if (k < sz - 1) { sum += zp[b]; cnt += 1.0f; }
            const float num = uo[b] + r * sum;
            const float den = 1.0f + r * cnt;
            u_out[(size_t)ii * di + base] = num / den;
        }
        col[0] = col[VISC_BATCH];
        col[1] = col[VISC_BATCH + 1];
    }
}

__global__ __launch_bounds__(GF_WG) void k_jacobi_visc3(ViscArgs a) {
    const int c = blockIdx.z % 3, chunk = blockIdx.z / 3;
    // the box's faces of component c, and the full face array's strides
    const int sx = a.bx + (c == 0), sy = a.by + (c == 1), sz = a.bz + (c == 2);
    const size_t dj = (size_t)(a.nz + (c == 2));
    const size_t di = (size_t)(a.ny + (c == 1)) * dj;
    const size_t org = (size_t)a.ox * di + (size_t)a.oy * dj + (size_t)a.oz;
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i0 = chunk * VISC_CHUNK;
    if (k >= sz || j >= sy || i0 >= sx) return;
    const int i1 = min(i0 + VISC_CHUNK, sx);
    if (c == 0)
        visc_column(a.old0 + org, a.in0 + org, a.out0 + org, sx, sy, sz, di, dj, i0, i1, j, k, a.r);
    else if (c == 1)
        visc_column(a.old1 + org, a.in1 + org, a.out1 + org, sx, sy, sz, di, dj, i0, i1, j, k, a.r);
    else
        visc_column(a.old2 + org, a.in2 + org, a.out2 + org, sx, sy, sz, di, dj, i0, i1, j, k, a.r);
}

// ---- per-particle attribute transfer (S2.15 colour, S2.18 scalar) --------
// P2G of `nc` channels per particle onto the cell-centred grid at the current
// positions (solver3d.py:562-605, 670-707), normalise (607-631, 710-727) and
// trilinear gather back (633-664, 729-755).  The scatter sums fixed-point
// integers with integer atomics, so the result does not depend on the order
// of arrival: each contribution is the reference's FP32 product value*w (and
// w), scaled by a power of two chosen from the largest |value| (values by
// 2^(48-E) with max |value| < 2^E, weights by 2^48), then rounded to an
// integer.  A cell sum stays below 2^63 while the particles of the 2x2x2
// cells around it number fewer than 32768; the rounding error is at most
// 2^-49 of max |value| per contribution, far below FP32 resolution.
// The cell sums are converted back to FP32 before the reference's FP32
// normalisation.  Attributes are stored in particle-id order; slot p of the
// cell-sorted arrays holds particle pid[p].

// Zero the fixed-point sums and the max-|value| word.
__global__ void k_attr_clear(int64_t* a, long long n, uint32_t* scratch) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
    if (i == 0) scratch[0] = 0;
}

// scratch[0] = bits of max |attr| (non-negative floats order as unsigned).
__global__ void k_attr_absmax(const float* attr, long long n, uint32_t* scratch) {
    uint32_t m = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const uint32_t b = __float_as_uint(attr[i]) & 0x7fffffffu;
        m = b > m && b < 0x7f800000u ? b : m;  // NaN/inf ignored for the scale
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(scratch, m);
}

// Fixed-point exponents: values get 2^(48 - E) with |value| < 2^E, weights 2^48.
__device__ __forceinline__ int attr_exp(const uint32_t* scratch) {
    int e;
    (void)frexpf(__uint_as_float(scratch[0]), &e);
    return 48 - e;
}

__device__ __forceinline__ void attr_cell_coords(const float* pos, int p, float dx, float& fx,
                                                 float& fy, float& fz) {
    fx = pos[3 * p + 0] / dx - 0.5f;
    fy = pos[3 * p + 1] / dx - 0.5f;
    fz = pos[3 * p + 2] / dx - 0.5f;
}

// A workgroup's 256 consecutive slots hold particles that were cell-sorted at
// the start of the substep and have moved by about a cell since, so their
// base cells fill a small box.  The group sums its contributions in LDS over
// that box (integer sums, so the order of the LDS atomics does not matter)
// and adds the non-zero cells to the grid sums; a group whose box exceeds the
// LDS budget adds each contribution to the grid directly.
constexpr int ATTR_LDS_BYTES = 48 * 1024;

template <int NC>
__global__ __launch_bounds__(GF_WG) void k_attr_scatter(const float* pos, const int32_t* pid,
                                                        const float* attr, int n, int nx, int ny,
                                                        int nz, float dx,
                                                        const uint32_t* scratch, int64_t* acc) {
    constexpr int CH = NC + 1;
    constexpr int CAP = ATTR_LDS_BYTES / (8 * CH);  // box cells
    __shared__ unsigned long long box[CAP * CH];
    __shared__ int red[6][GF_WG / 64];
    const int p = blockIdx.x * GF_WG + threadIdx.x;
    const double sv = ldexp(1.0, attr_exp(scratch)), sw = ldexp(1.0, 48);
    const long long cells = (long long)nx * ny * nz;
    bool ok = false;
    int i0 = 0, j0 = 0, k0 = 0;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    float v[NC];
    for (int c = 0; c < NC; ++c) v[c] = 0.0f;
    if (p < n) {
        float fx, fy, fz;
        attr_cell_coords(pos, p, dx, fx, fy, fz);
        // No corner is in range outside (-2, n+1) per axis; this also drops
        // NaN positions before the float->int conversion.
        ok = fx > -2.0f && fx < (float)(nx + 1) && fy > -2.0f && fy < (float)(ny + 1) &&
             fz > -2.0f && fz < (float)(nz + 1);
        if (ok) {
            i0 = (int)floorf(fx); j0 = (int)floorf(fy); k0 = (int)floorf(fz);
            sx = fx - (float)i0; sy = fy - (float)j0; sz = fz - (float)k0;
            const int id = pid ? pid[p] : p;
            for (int c = 0; c < NC; ++c) v[c] = attr[(size_t)id * NC + c];
        }
    }
    // bounding box of the group's base cells
    int b[6] = {ok ? i0 : INT_MAX, ok ? j0 : INT_MAX, ok ? k0 : INT_MAX,
                ok ? -i0 : INT_MAX, ok ? -j0 : INT_MAX, ok ? -k0 : INT_MAX};
    for (int q = 0; q < 6; ++q)
        for (int off = 32; off > 0; off >>= 1) b[q] = min(b[q], __shfl_xor(b[q], off, 64));
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 6; ++q) red[q][wave] = b[q];
    __syncthreads();
    for (int q = 0; q < 6; ++q) {
        int m = red[q][0];
        for (int w = 1; w < GF_WG / 64; ++w) m = min(m, red[q][w]);
        b[q] = m;
    }
    if (b[0] == INT_MAX) return;  // no particle of the group is near the grid
    const int lo_i = b[0], lo_j = b[1], lo_k = b[2];
    const int by = -b[4] - lo_j + 2, bz = -b[5] - lo_k + 2;
    const long long nbox = (long long)(-b[3] - lo_i + 2) * by * bz;
    const bool in_lds = nbox <= CAP;  // uniform over the group
    const int nb = in_lds ? (int)nbox : 0;
    if (in_lds) {
        for (int e = threadIdx.x; e < nb * CH; e += GF_WG) box[e] = 0;
        __syncthreads();
    }
    if (ok) {
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                    if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
                    const float wx = di ? sx : 1.0f - sx;
                    const float wy = dj ? sy : 1.0f - sy;
                    const float wz = dk ? sz : 1.0f - sz;
                    const float w = wx * wy * wz;
                    unsigned long long q[CH];
                    for (int c = 0; c < NC; ++c)
                        q[c] = (unsigned long long)llrint((double)(v[c] * w) * sv);
                    q[NC] = (unsigned long long)llrint((double)w * sw);
                    if (in_lds) {
                        const int e = ((ii - lo_i) * by + (jj - lo_j)) * bz + (kk - lo_k);
                        for (int c = 0; c < CH; ++c) atomicAdd(&box[c * nb + e], q[c]);
                    } else {
                        const long long cell = ((long long)ii * ny + jj) * nz + kk;
                        for (int c = 0; c < CH; ++c)
                            atomicAdd((unsigned long long*)(acc + c * cells + cell), q[c]);
                    }
                }
    }
    if (!in_lds) return;
    __syncthreads();
    // A cell with no weight has no value either: |value*w| * 2^(48-E) is
    // below w * 2^48, so both round to 0 together.
    const int byz = by * bz;
    for (int e = threadIdx.x; e < nb; e += GF_WG) {
        if (box[NC * nb + e] == 0) continue;
        const int ei = e / byz, r = e - ei * byz, ej = r / bz, ek = r - ej * bz;
        const long long cell = ((long long)(lo_i + ei) * ny + (lo_j + ej)) * nz + (lo_k + ek);
        for (int c = 0; c < CH; ++c) {
            const unsigned long long x = box[c * nb + e];
            if (x) atomicAdd((unsigned long long*)(acc + c * cells + cell), x);
        }
    }
}

// grid[c] = normalised channel c, grid[nc] = weight sum (FP32).  mode 0: the
// colour rule value * (1/w); mode 1: the scalar rule value / w; 0 where
// w <= 1e-8.
__global__ void k_attr_normalize(const int64_t* acc, int nc, long long cells,
                                 const uint32_t* scratch, int mode, float* grid) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const double iv = ldexp(1.0, -attr_exp(scratch)), iw = ldexp(1.0, -48);
    const float w = (float)((double)acc[nc * cells + i] * iw);
    const float inv = 1.0f / w;
    for (int c = 0; c < nc; ++c) {
        const float v = (float)((double)acc[c * cells + i] * iv);
        grid[c * cells + i] = w > 1.0e-8f ? (mode == 0 ? v * inv : v / w) : 0.0f;
    }
    grid[nc * cells + i] = w;
}

// The block forms' grids (S2.15.1 / S2.18.1 semantics): each channel's exact
// sum, rounded to FP32, is added to the caller's FP32 grid g[c] (g[nc] = the
// weight grid), as the reference's atomics add into it.
struct AttrGrids {
    float* g[4];
};

__global__ void k_attr_accumulate(const int64_t* acc, int nc, long long cells,
                                  const uint32_t* scratch, AttrGrids out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const double iv = ldexp(1.0, -attr_exp(scratch)), iw = ldexp(1.0, -48);
    for (int c = 0; c <= nc; ++c) {
        const int64_t a = acc[c * cells + i];
        if (a == 0) continue;
        out.g[c][i] = out.g[c][i] + (float)((double)a * (c == nc ? iw : iv));
    }
}

// k3_normalize_color (solver3d.py:607-631) / k3_normalize_scalar (710-727).
__global__ void k_attr_normalize_grids(int nc, long long cells, int mode, AttrGrids g) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const float w = g.g[nc][i];
    if (w > 1.0e-8f) {
        const float inv = 1.0f / w;
        for (int c = 0; c < nc; ++c) g.g[c][i] = mode == 0 ? g.g[c][i] * inv : g.g[c][i] / w;
    } else {
        for (int c = 0; c < nc; ++c) g.g[c][i] = 0.0f;
    }
}

// k3_g2p_color (solver3d.py:633-664) / k3_g2p_scalar (729-755) on the grids.
__global__ void k_attr_gather_grids(const float* pos, float* attr, int nc, int n, int nx, int ny,
                                    int nz, float dx, AttrGrids g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float fx, fy, fz;
    attr_cell_coords(pos, p, dx, fx, fy, fz);
    const float w = sample3(g.g[nc], fx, fy, fz, nx, ny, nz);
    if (!(w > 1.0e-6f)) return;
    for (int c = 0; c < nc; ++c) attr[(size_t)p * nc + c] = sample3(g.g[c], fx, fy, fz, nx, ny, nz);
}

__global__ void k_attr_gather(const float* pos, const int32_t* pid, float* attr, int nc, int n,
                              int nx, int ny, int nz, float dx, const float* grid) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long long cells = (long long)nx * ny * nz;
    float fx, fy, fz;
    attr_cell_coords(pos, p, dx, fx, fy, fz);
    const float w = sample3(grid + nc * cells, fx, fy, fz, nx, ny, nz);
    if (!(w > 1.0e-6f)) return;
    const int id = pid[p];
    for (int c = 0; c < nc; ++c)
        attr[(size_t)id * nc + c] = sample3(grid + c * cells, fx, fy, fz, nx, ny, nz);
}

}  // namespace

#define GRID_ARGS_OK(who)                                                         \
    GF_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, who ": grid %dx%dx%d", nx, ny, nz)

extern "C" int gf_k3_clear_grid(float* u, float* v, float* w, float* uw, float* vw, float* ww,
                                int32_t* marker, int nx, int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_clear_grid");
    GF_REQUIRE(u && v && w && uw && vw && ww && marker, "gf_k3_clear_grid: null buffer");
    hipLaunchKernelGGL(k_clear_grid, grid_for((long long)(nx + 1) * (ny + 1) * (nz + 1)),
                       dim3(GF_WG), 0, (hipStream_t)stream, u, v, w, uw, vw, ww, marker, nx, ny, nz);
    GF_LAUNCH_CHECK("k_clear_grid");
    return 0;
}

extern "C" int gf_k3_mark_fluid(const float* pos, int n, int32_t* marker, float dx, int nx,
                                int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_mark_fluid");
    GF_REQUIRE(n >= 0 && (n == 0 || pos) && marker, "gf_k3_mark_fluid: bad buffers");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_mark_fluid, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, pos, n,
                       marker, dx, nx, ny, nz);
    GF_LAUNCH_CHECK("k_mark_fluid");
    return 0;
}

extern "C" int gf_k3_normalize(float* u, float* v, float* w, const float* uw, const float* vw,
                               const float* ww, float* us, float* vs, float* ws, int nx, int ny,
                               int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_normalize");
    GF_REQUIRE(u && v && w && uw && vw && ww && us && vs && ws, "gf_k3_normalize: null buffer");
    hipLaunchKernelGGL(k_normalize, grid_for((long long)(nx + 1) * (ny + 1) * (nz + 1)),
                       dim3(GF_WG), 0, (hipStream_t)stream, u, v, w, uw, vw, ww, us, vs, ws, nx,
                       ny, nz);
    GF_LAUNCH_CHECK("k_normalize");
    return 0;
}

extern "C" int gf_k3_add_gravity(float* v, int nx, int ny, int nz, float g, float dt,
                                 void* stream) {
    GRID_ARGS_OK("gf_k3_add_gravity");
    GF_REQUIRE(v, "gf_k3_add_gravity: null buffer");
    long long n = (long long)nx * (ny + 1) * nz;
    hipLaunchKernelGGL(k_add_gravity, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, v, n, g, dt);
    GF_LAUNCH_CHECK("k_add_gravity");
    return 0;
}

extern "C" int gf_k3_enforce_solid_bc(float* u, float* v, float* w, const int32_t* marker,
                                      const float* solid_u, const float* solid_v,
                                      const float* solid_w, int nx, int ny, int nz,
                                      void* stream) {
    GRID_ARGS_OK("gf_k3_enforce_solid_bc");
    GF_REQUIRE(u && v && w && marker, "gf_k3_enforce_solid_bc: null buffer");
    GF_REQUIRE(nx + 1 <= 65535 && ny + 1 <= 4 * 65535, "gf_k3_enforce_solid_bc: grid %d x %d",
               nx, ny);
    const dim3 grid((unsigned)((nz + 1 + 63) / 64), (unsigned)((ny + 1 + 3) / 4),
                    (unsigned)(nx + 1));
    hipLaunchKernelGGL(k_enforce_solid_bc, grid, dim3(GF_WG), 0, (hipStream_t)stream, u, v, w,
                       marker, solid_u, solid_v, solid_w, nx, ny, nz);
    GF_LAUNCH_CHECK("k_enforce_solid_bc");
    return 0;
}

extern "C" int gf_k3_compute_divergence(const float* u, const float* v, const float* w,
                                        float* div, const int32_t* marker, int nx, int ny, int nz,
                                        float dx, float dt, float rho, void* stream) {
    GRID_ARGS_OK("gf_k3_compute_divergence");
    GF_REQUIRE(u && v && w && div && marker, "gf_k3_compute_divergence: null buffer");
    hipLaunchKernelGGL(k_divergence, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, u, v, w, div, marker, nx, ny, nz, dx, dt, rho);
    GF_LAUNCH_CHECK("k_divergence");
    return 0;
}

extern "C" int gf_k3_jacobi_pressure(const float* p_in, float* p_out, const float* div,
                                     const int32_t* marker, int nx, int ny, int nz,
                                     void* stream) {
    GRID_ARGS_OK("gf_k3_jacobi_pressure");
    GF_REQUIRE(p_in && p_out && div && marker && p_in != p_out,
               "gf_k3_jacobi_pressure: bad buffers");
    hipLaunchKernelGGL(k_jacobi_dense, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, p_in, p_out, div, marker, nx, ny, nz, Dom{},
                       BlkMask{});
    GF_LAUNCH_CHECK("k_jacobi_dense");
    return 0;
}

extern "C" int gf_k3_gauss_seidel_rb(float* p, const float* div, const int32_t* marker,
                                     int nx, int ny, int nz, int color, void* stream) {
    GRID_ARGS_OK("gf_k3_gauss_seidel_rb");
    GF_REQUIRE(p && div && marker, "gf_k3_gauss_seidel_rb: null buffer");
    GF_REQUIRE(color == 0 || color == 1, "gf_k3_gauss_seidel_rb: color %d", color);
    hipLaunchKernelGGL(k_gsrb_dense, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, p, div, marker, nx, ny, nz, color, Dom{});
    GF_LAUNCH_CHECK("k_gsrb_dense");
    return 0;
}

extern "C" int gf_k3_jacobi_visc(const float* u_old, const float* u_in, float* u_out, int sx,
                                 int sy, int sz, float r, void* stream) {
    GF_REQUIRE(sx > 0 && sy > 0 && sz > 0, "gf_k3_jacobi_visc: shape %d x %d x %d", sx, sy, sz);
    GF_REQUIRE(sy <= 4 * 65535 && sx <= 65535, "gf_k3_jacobi_visc: shape %d x %d x %d", sx, sy,
               sz);
    GF_REQUIRE(u_old && u_in && u_out && u_in != u_out && u_old != u_out,
               "gf_k3_jacobi_visc: bad buffers");
    const dim3 grid((unsigned)((sz + 63) / 64), (unsigned)((sy + 3) / 4), (unsigned)sx);
    hipLaunchKernelGGL(k_jacobi_visc, grid, dim3(GF_WG), 0, (hipStream_t)stream, u_old, u_in,
                       u_out, sx, sy, sz, r);
    GF_LAUNCH_CHECK("k_jacobi_visc");
    return 0;
}

namespace {
int visc3_launch(const float* old_u, const float* old_v, const float* old_w, const float* in_u,
                 const float* in_v, const float* in_w, float* out_u, float* out_v, float* out_w,
                 int nx, int ny, int nz, int ox, int oy, int oz, int bx, int by, int bz, float r,
                 void* stream, const char* who) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "%s: grid %d x %d x %d", who, nx, ny, nz);
    GF_REQUIRE(bx > 0 && by > 0 && bz > 0 && ox >= 0 && oy >= 0 && oz >= 0 && ox + bx <= nx &&
                   oy + by <= ny && oz + bz <= nz,
               "%s: box origin (%d,%d,%d) extent (%d,%d,%d) not inside the %d x %d x %d grid",
               who, ox, oy, oz, bx, by, bz, nx, ny, nz);
    GF_REQUIRE(by + 1 <= 4 * 65535, "%s: box y extent %d", who, by);
    GF_REQUIRE(old_u && old_v && old_w && in_u && in_v && in_w && out_u && out_v && out_w,
               "%s: null buffer", who);
    GF_REQUIRE(out_u != in_u && out_v != in_v && out_w != in_w && out_u != old_u &&
                   out_v != old_v && out_w != old_w,
               "%s: out must not alias in or old", who);
    ViscArgs a;
    a.old0 = old_u; a.old1 = old_v; a.old2 = old_w;
    a.in0 = in_u; a.in1 = in_v; a.in2 = in_w;
    a.out0 = out_u; a.out1 = out_v; a.out2 = out_w;
    a.nx = nx; a.ny = ny; a.nz = nz; a.r = r;
    a.ox = ox; a.oy = oy; a.oz = oz;
    a.bx = bx; a.by = by; a.bz = bz;
    a.nchunk = (bx + 1 + VISC_CHUNK - 1) / VISC_CHUNK;
    GF_REQUIRE(3LL * a.nchunk <= 65535, "%s: box x extent %d", who, bx);
    const dim3 grid((unsigned)((bz + 1 + 63) / 64), (unsigned)((by + 1 + 3) / 4),
                    (unsigned)(3 * a.nchunk));
    hipLaunchKernelGGL(k_jacobi_visc3, grid, dim3(GF_WG), 0, (hipStream_t)stream, a);
    GF_LAUNCH_CHECK("k_jacobi_visc3");
    return 0;
}
}  // namespace

extern "C" int gf_k3_jacobi_visc3(const float* old_u, const float* old_v, const float* old_w,
                                  const float* in_u, const float* in_v, const float* in_w,
                                  float* out_u, float* out_v, float* out_w, int nx, int ny,
                                  int nz, float r, void* stream) {
    return visc3_launch(old_u, old_v, old_w, in_u, in_v, in_w, out_u, out_v, out_w, nx, ny, nz,
                        0, 0, 0, nx, ny, nz, r, stream, "gf_k3_jacobi_visc3");
}

extern "C" int gf_k3_jacobi_visc3_box(const float* old_u, const float* old_v, const float* old_w,
                                      const float* in_u, const float* in_v, const float* in_w,
                                      float* out_u, float* out_v, float* out_w, int nx, int ny,
                                      int nz, int ox, int oy, int oz, int bx, int by, int bz,
                                      float r, void* stream) {
    return visc3_launch(old_u, old_v, old_w, in_u, in_v, in_w, out_u, out_v, out_w, nx, ny, nz,
                        ox, oy, oz, bx, by, bz, r, stream, "gf_k3_jacobi_visc3_box");
}

extern "C" int gf_k3_apply_A(const float* x, float* y, const int32_t* marker,
                             const int32_t* done, int nx, int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_apply_A");
    GF_REQUIRE(x && y && marker && x != y, "gf_k3_apply_A: bad buffers");
    hipLaunchKernelGGL(k_apply_A_dense, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, x, y, marker, done, nx, ny, nz, Dom{});
    GF_LAUNCH_CHECK("k_apply_A_dense");
    return 0;
}

extern "C" int gf_k3_compute_diag(const int32_t* marker, float* diag, int nx, int ny, int nz,
                                  void* stream) {
    GRID_ARGS_OK("gf_k3_compute_diag");
    GF_REQUIRE(marker && diag, "gf_k3_compute_diag: null buffer");
    hipLaunchKernelGGL(k_compute_diag_dense, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, marker, diag, nx, ny, nz);
    GF_LAUNCH_CHECK("k_compute_diag_dense");
    return 0;
}

extern "C" int gf_k3_apply_invM(const float* r, const float* diag, float* z,
                                const int32_t* marker, const int32_t* done, int nx, int ny,
                                int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_apply_invM");
    GF_REQUIRE(r && diag && z && marker, "gf_k3_apply_invM: null buffer");
    long long n = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_apply_invM_dense, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, r,
                       diag, z, marker, done, nx, ny, nz, Dom{});
    GF_LAUNCH_CHECK("k_apply_invM_dense");
    return 0;
}

extern "C" int gf_k3_dot_fluid(const float* a, const float* b, float* out, const int32_t* marker,
                               const int32_t* done, int nx, int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_dot_fluid");
    GF_REQUIRE(a && b && out && marker, "gf_k3_dot_fluid: null buffer");
    long long n = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_dot_dense, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, a, b, out,
                       marker, done, nx, ny, nz, Dom{});
    GF_LAUNCH_CHECK("k_dot_dense");
    return 0;
}

extern "C" int gf_k3_axpy_devscalar(const float* a_dev, float sign, const float* x,
                                    const float* y_in, float* y_out, const int32_t* marker,
                                    const int32_t* done, int nx, int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_axpy_devscalar");
    GF_REQUIRE(a_dev && x && y_in && y_out && marker, "gf_k3_axpy_devscalar: null buffer");
    long long n = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_axpy_devscalar_dense, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream,
                       a_dev, sign, x, y_in, y_out, marker, done, nx, ny, nz, Dom{});
    GF_LAUNCH_CHECK("k_axpy_devscalar_dense");
    return 0;
}

// ---- block-sparse and per-tile forms (S2.16, S2.6.4-S2.6.6) ----------------

#define BLOCKS_COVER_OK(who)                                                           \
    GF_REQUIRE(bs >= 1 && bs <= 64 && (long long)nbx * bs >= nx &&                     \
                   (long long)nby * bs >= ny && (long long)nbz * bs >= nz,              \
               who ": %d^3 blocks %dx%dx%d do not cover the grid", bs, nbx, nby, nbz)

#define TILES_OK(who)                                                                  \
    GF_REQUIRE(coords && n_active && bs >= 1 && bs <= 64 && max_blocks >= 0 &&         \
                   (long long)max_blocks * bs * bs * bs <= (long long)INT32_MAX,                  \
               who ": tile list (%d blocks of %d^3)", max_blocks, bs)

namespace {
inline Dom tiles(const int32_t* coords, const int32_t* n_active, int bs) {
    Dom d;
    d.coords = coords; d.n_active = n_active; d.bs = bs;
    return d;
}
inline dim3 tile_grid(int max_blocks, int bs) {
    return grid_for((long long)max_blocks * bs * bs * bs);
}
}  // namespace

extern "C" int gf_k3_mark_active_blocks(const int32_t* marker, int32_t* block_active, int nx,
                                        int ny, int nz, int bs, int nbx, int nby, int nbz,
                                        void* stream) {
    GRID_ARGS_OK("gf_k3_mark_active_blocks");
    BLOCKS_COVER_OK("gf_k3_mark_active_blocks");
    GF_REQUIRE(marker && block_active, "gf_k3_mark_active_blocks: null buffer");
    hipLaunchKernelGGL(k_mark_blocks, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, marker, block_active, nx, ny, nz, bs, nby, nbz);
    GF_LAUNCH_CHECK("k_mark_blocks");
    return 0;
}

extern "C" int gf_k3_jacobi_pressure_blocksparse(const float* p_in, float* p_out, const float* div,
                                                 const int32_t* marker,
                                                 const int32_t* block_active, int nx, int ny,
                                                 int nz, int bs, int nbx, int nby, int nbz,
                                                 void* stream) {
    GRID_ARGS_OK("gf_k3_jacobi_pressure_blocksparse");
    BLOCKS_COVER_OK("gf_k3_jacobi_pressure_blocksparse");
    GF_REQUIRE(p_in && p_out && div && marker && block_active && p_in != p_out,
               "gf_k3_jacobi_pressure_blocksparse: bad buffers");
    BlkMask bm;
    bm.active = block_active; bm.bs = bs; bm.nby = nby; bm.nbz = nbz;
    hipLaunchKernelGGL(k_jacobi_dense, grid_for((long long)nx * ny * nz), dim3(GF_WG), 0,
                       (hipStream_t)stream, p_in, p_out, div, marker, nx, ny, nz, Dom{}, bm);
    GF_LAUNCH_CHECK("k_jacobi_dense (block-sparse)");
    return 0;
}

extern "C" int gf_k3_jacobi_pressure_per_tile(const float* p_in, float* p_out, const float* div,
                                              const int32_t* marker, const int32_t* coords,
                                              int max_blocks, int bs, const int32_t* n_active,
                                              int nx, int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_jacobi_pressure_per_tile");
    TILES_OK("gf_k3_jacobi_pressure_per_tile");
    GF_REQUIRE(p_in && p_out && div && marker && p_in != p_out,
               "gf_k3_jacobi_pressure_per_tile: bad buffers");
    if (max_blocks == 0) return 0;
    hipLaunchKernelGGL(k_jacobi_dense, tile_grid(max_blocks, bs), dim3(GF_WG), 0,
                       (hipStream_t)stream, p_in, p_out, div, marker, nx, ny, nz,
                       tiles(coords, n_active, bs), BlkMask{});
    GF_LAUNCH_CHECK("k_jacobi_dense (per tile)");
    return 0;
}

extern "C" int gf_k3_gauss_seidel_rb_per_tile(float* p, const float* div, const int32_t* marker,
                                              const int32_t* coords, int max_blocks, int bs,
                                              int color, const int32_t* n_active, int nx, int ny,
                                              int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_gauss_seidel_rb_per_tile");
    TILES_OK("gf_k3_gauss_seidel_rb_per_tile");
    GF_REQUIRE(p && div && marker, "gf_k3_gauss_seidel_rb_per_tile: null buffer");
    GF_REQUIRE(color == 0 || color == 1, "gf_k3_gauss_seidel_rb_per_tile: color %d", color);
    if (max_blocks == 0) return 0;
    hipLaunchKernelGGL(k_gsrb_dense, tile_grid(max_blocks, bs), dim3(GF_WG), 0,
                       (hipStream_t)stream, p, div, marker, nx, ny, nz, color,
                       tiles(coords, n_active, bs));
    GF_LAUNCH_CHECK("k_gsrb_dense (per tile)");
    return 0;
}

extern "C" int gf_k3_apply_A_per_tile(const float* x, float* y, const int32_t* marker,
                                      const int32_t* coords, int max_blocks, int bs,
                                      const int32_t* n_active, const int32_t* done, int nx,
                                      int ny, int nz, void* stream) {
    GRID_ARGS_OK("gf_k3_apply_A_per_tile");
    TILES_OK("gf_k3_apply_A_per_tile");
    GF_REQUIRE(x && y && marker && x != y, "gf_k3_apply_A_per_tile: bad buffers");
    if (max_blocks == 0) return 0;
    hipLaunchKernelGGL(k_apply_A_dense, tile_grid(max_blocks, bs), dim3(GF_WG), 0,
                       (hipStream_t)stream, x, y, marker, done, nx, ny, nz,
                       tiles(coords, n_active, bs));
    GF_LAUNCH_CHECK("k_apply_A_dense (per tile)");
    return 0;
}

extern "C" int gf_k3_apply_invM_per_tile(const float* r, const float* diag, float* z,
                                         const int32_t* marker, const int32_t* coords,
                                         int max_blocks, int bs, const int32_t* n_active,
                                         const int32_t* done, int nx, int ny, int nz,
                                         void* stream) {
    GRID_ARGS_OK("gf_k3_apply_invM_per_tile");
    TILES_OK("gf_k3_apply_invM_per_tile");
    GF_REQUIRE(r && diag && z && marker, "gf_k3_apply_invM_per_tile: null buffer");
    if (max_blocks == 0) return 0;
    hipLaunchKernelGGL(k_apply_invM_dense, tile_grid(max_blocks, bs), dim3(GF_WG), 0,
                       (hipStream_t)stream, r, diag, z, marker, done, nx, ny, nz,
                       tiles(coords, n_active, bs));
    GF_LAUNCH_CHECK("k_apply_invM_dense (per tile)");
    return 0;
}

extern "C" int gf_k3_dot_fluid_per_tile(const float* a, const float* b, float* out,
                                        const int32_t* marker, const int32_t* coords,
                                        int max_blocks, int bs, const int32_t* n_active,
                                        const int32_t* done, int nx, int ny, int nz,
                                        void* stream) {
    GRID_ARGS_OK("gf_k3_dot_fluid_per_tile");
    TILES_OK("gf_k3_dot_fluid_per_tile");
    GF_REQUIRE(a && b && out && marker, "gf_k3_dot_fluid_per_tile: null buffer");
    if (max_blocks == 0) return 0;
    hipLaunchKernelGGL(k_dot_dense, tile_grid(max_blocks, bs), dim3(GF_WG), 0,
                       (hipStream_t)stream, a, b, out, marker, done, nx, ny, nz,
                       tiles(coords, n_active, bs));
    GF_LAUNCH_CHECK("k_dot_dense (per tile)");
    return 0;
}

extern "C" int gf_k3_axpy_devscalar_per_tile(const float* a_dev, float sign, const float* x,
                                             const float* y_in, float* y_out,
                                             const int32_t* marker, const int32_t* coords,
                                             int max_blocks, int bs, const int32_t* n_active,
                                             const int32_t* done, int nx, int ny, int nz,
                                             void* stream) {
    GRID_ARGS_OK("gf_k3_axpy_devscalar_per_tile");
    TILES_OK("gf_k3_axpy_devscalar_per_tile");
    GF_REQUIRE(a_dev && x && y_in && y_out && marker,
               "gf_k3_axpy_devscalar_per_tile: null buffer");
    if (max_blocks == 0) return 0;
    hipLaunchKernelGGL(k_axpy_devscalar_dense, tile_grid(max_blocks, bs), dim3(GF_WG), 0,
                       (hipStream_t)stream, a_dev, sign, x, y_in, y_out, marker, done, nx, ny,
                       nz, tiles(coords, n_active, bs));
    GF_LAUNCH_CHECK("k_axpy_devscalar_dense (per tile)");
    return 0;
}

extern "C" int gf_k3_scalar_op(int op, float* a, float* b, float* c, float tol, int32_t* done,
                               void* stream) {
    GF_REQUIRE(op >= 0 && op <= 5, "gf_k3_scalar_op: op %d", op);
    const bool need_a = op <= 4, need_b = op == 1 || op == 4, need_c = op == 1 || op == 2 || op == 3;
    const bool need_done = op == 4 || op == 5;
    GF_REQUIRE((!need_a || a) && (!need_b || b) && (!need_c || c) && (!need_done || done),
               "gf_k3_scalar_op: op %d missing a buffer", op);
    hipLaunchKernelGGL(k_scalar_op, dim3(1), dim3(1), 0, (hipStream_t)stream, op, a, b, c, tol, done);
    GF_LAUNCH_CHECK("k_scalar_op");
    return 0;
}

extern "C" int gf_k3_subtract_pressure_grad(float* u, float* v, float* w, const float* p,
                                            const int32_t* marker, int nx, int ny, int nz,
                                            float dx, float dt, float rho, void* stream) {
    GRID_ARGS_OK("gf_k3_subtract_pressure_grad");
    GF_REQUIRE(u && v && w && p && marker, "gf_k3_subtract_pressure_grad: null buffer");
    hipLaunchKernelGGL(k_subtract_grad, grid_for((long long)(nx + 1) * (ny + 1) * (nz + 1)),
                       dim3(GF_WG), 0, (hipStream_t)stream, u, v, w, p, marker, nx, ny, nz, dx,
                       dt, rho);
    GF_LAUNCH_CHECK("k_subtract_grad");
    return 0;
}

extern "C" int gf_k3_g2p_and_advect(float* pos, float* vel, const float* u, const float* v,
                                    const float* w, const float* us, const float* vs,
                                    const float* ws, int n, float dx, float dt, int nx, int ny,
                                    int nz, float flip_blend, const float* dom, void* stream) {
    GRID_ARGS_OK("gf_k3_g2p_and_advect");
    GF_REQUIRE(n >= 0 && dom, "gf_k3_g2p_and_advect: bad arguments");
    if (n == 0) return 0;
    GF_REQUIRE(pos && vel && u && v && w && us && vs && ws, "gf_k3_g2p_and_advect: null buffer");
    hipLaunchKernelGGL(k_g2p_advect_dense, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, pos,
                       vel, u, v, w, us, vs, ws, n, dx, dt, nx, ny, nz, flip_blend, dom[0],
                       dom[1], dom[2]);
    GF_LAUNCH_CHECK("k_g2p_advect_dense");
    return 0;
}

extern "C" int gf_k3_g2p_apic_advect(float* pos, float* vel, float* affine_c, const float* u,
                                     const float* v, const float* w, int n, float dx, float dt,
                                     int nx, int ny, int nz, const float* dom, void* stream) {
    GRID_ARGS_OK("gf_k3_g2p_apic_advect");
    GF_REQUIRE(n >= 0 && dom, "gf_k3_g2p_apic_advect: bad arguments");
    if (n == 0) return 0;
    GF_REQUIRE(pos && vel && affine_c && u && v && w, "gf_k3_g2p_apic_advect: null buffer");
    hipLaunchKernelGGL(k_g2p_apic_dense, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, pos,
                       vel, affine_c, u, v, w, n, dx, dt, nx, ny, nz, dom[0], dom[1], dom[2]);
    GF_LAUNCH_CHECK("k_g2p_apic_dense");
    return 0;
}

extern "C" int gf_k3_vmax_sq_reduce(const float* vel, int n, float* out, void* stream) {
    GF_REQUIRE(n >= 0 && out, "gf_k3_vmax_sq_reduce: bad arguments");
    if (n == 0) return 0;
    GF_REQUIRE(vel, "gf_k3_vmax_sq_reduce: null buffer");
    hipLaunchKernelGGL(k_vmax_sq, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, vel, n, out);
    GF_LAUNCH_CHECK("k_vmax_sq");
    return 0;
}

extern "C" int gf_scene_attr_transfer(const gf_scene_t* s, float* attr, int channels, int mode,
                                      int64_t* acc, float* grid, uint32_t* scratch,
                                      void* stream) {
    GF_REQUIRE(s, "gf_scene_attr_transfer: null scene");
    GF_REQUIRE(s->n_scenes <= 1, "gf_scene_attr_transfer: not for packed scenes");
    GF_REQUIRE(channels >= 1 && channels <= 3, "gf_scene_attr_transfer: channels %d", channels);
    GF_REQUIRE(mode == GF_ATTR_COLOR || mode == GF_ATTR_SCALAR, "gf_scene_attr_transfer: mode %d",
               mode);
    GF_REQUIRE(s->nx > 0 && s->ny > 0 && s->nz > 0 && s->dx > 0.0f,
               "gf_scene_attr_transfer: grid");
    if (s->n == 0) return 0;
    GF_REQUIRE(attr && acc && grid && scratch && s->pos && s->pid,
               "gf_scene_attr_transfer: null buffer");
    hipStream_t st = (hipStream_t)stream;
    const long long cells = (long long)s->nx * s->ny * s->nz;
    const long long nacc = (long long)(channels + 1) * cells;
    hipLaunchKernelGGL(k_attr_clear, grid_for(nacc), dim3(GF_WG), 0, st, acc, nacc, scratch);
    const long long nv = (long long)s->n * channels;
    const unsigned gb = (unsigned)std::min<long long>((nv + GF_WG - 1) / GF_WG, 2048);
    hipLaunchKernelGGL(k_attr_absmax, dim3(gb), dim3(GF_WG), 0, st, attr, nv, scratch);
    auto scatter = channels == 1 ? k_attr_scatter<1>
                   : channels == 2 ? k_attr_scatter<2> : k_attr_scatter<3>;
    hipLaunchKernelGGL(scatter, grid_for(s->n), dim3(GF_WG), 0, st, s->pos, s->pid, attr, s->n,
                       s->nx, s->ny, s->nz, s->dx, scratch, acc);
    hipLaunchKernelGGL(k_attr_normalize, grid_for(cells), dim3(GF_WG), 0, st, acc, channels, cells,
                       scratch, mode, grid);
    hipLaunchKernelGGL(k_attr_gather, grid_for(s->n), dim3(GF_WG), 0, st, s->pos, s->pid, attr,
                       channels, s->n, s->nx, s->ny, s->nz, s->dx, grid);
    GF_LAUNCH_CHECK("attribute transfer");
    return 0;
}


// ---- colour / scalar transfer one block at a time (S2.15.1-3, S2.18.1-3) ----

namespace {
int attr_block_ok(int channels, const float* g0, const float* g1, const float* g2,
                  const float* gw, int nx, int ny, int nz, const char* who) {
    GF_REQUIRE(channels == 1 || channels == 3, "%s: channels %d", who, channels);
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "%s: grid", who);
    GF_REQUIRE(g0 && gw && (channels == 1 || (g1 && g2)), "%s: null grid", who);
    return 0;
}
AttrGrids grids_of(int channels, float* g0, float* g1, float* g2, float* gw) {
    AttrGrids g;
    g.g[0] = g0; g.g[1] = g1; g.g[2] = g2; g.g[3] = nullptr;
    g.g[channels] = gw;
    return g;
}
}  // namespace

extern "C" int gf_k3_p2g_attr(const float* pos, const float* attr, int channels, int n,
                              float* g0, float* g1, float* g2, float* gw, int nx, int ny, int nz,
                              float dx, int64_t* acc, uint32_t* scratch, void* stream) {
    int rc = attr_block_ok(channels, g0, g1, g2, gw, nx, ny, nz, "gf_k3_p2g_attr");
    if (rc) return rc;
    GF_REQUIRE(n >= 0 && dx > 0.0f, "gf_k3_p2g_attr: n %d, dx", n);
    if (n == 0) return 0;
    GF_REQUIRE(pos && attr && acc && scratch, "gf_k3_p2g_attr: null buffer");
    hipStream_t st = (hipStream_t)stream;
    const long long cells = (long long)nx * ny * nz;
    const long long nacc = (long long)(channels + 1) * cells;
    hipLaunchKernelGGL(k_attr_clear, grid_for(nacc), dim3(GF_WG), 0, st, acc, nacc, scratch);
    const long long nv = (long long)n * channels;
    const unsigned gb = (unsigned)std::min<long long>((nv + GF_WG - 1) / GF_WG, 2048);
    hipLaunchKernelGGL(k_attr_absmax, dim3(gb), dim3(GF_WG), 0, st, attr, nv, scratch);
    auto scatter = channels == 1 ? k_attr_scatter<1> : k_attr_scatter<3>;
    hipLaunchKernelGGL(scatter, grid_for(n), dim3(GF_WG), 0, st, pos, nullptr, attr, n, nx, ny,
                       nz, dx, scratch, acc);
    hipLaunchKernelGGL(k_attr_accumulate, grid_for(cells), dim3(GF_WG), 0, st, acc, channels,
                       cells, scratch, grids_of(channels, g0, g1, g2, gw));
    GF_LAUNCH_CHECK("gf_k3_p2g_attr");
    return 0;
}

extern "C" int gf_k3_normalize_attr(int mode, int channels, float* g0, float* g1, float* g2,
                                    float* gw, int nx, int ny, int nz, void* stream) {
    int rc = attr_block_ok(channels, g0, g1, g2, gw, nx, ny, nz, "gf_k3_normalize_attr");
    if (rc) return rc;
    GF_REQUIRE(mode == GF_ATTR_COLOR || mode == GF_ATTR_SCALAR, "gf_k3_normalize_attr: mode %d",
               mode);
    const long long cells = (long long)nx * ny * nz;
    hipLaunchKernelGGL(k_attr_normalize_grids, grid_for(cells), dim3(GF_WG), 0,
                       (hipStream_t)stream, channels, cells, mode,
                       grids_of(channels, g0, g1, g2, gw));
    GF_LAUNCH_CHECK("k_attr_normalize_grids");
    return 0;
}

extern "C" int gf_k3_g2p_attr(const float* pos, float* attr, int channels, int n, const float* g0,
                              const float* g1, const float* g2, const float* gw, int nx, int ny,
                              int nz, float dx, void* stream) {
    int rc = attr_block_ok(channels, g0, g1, g2, gw, nx, ny, nz, "gf_k3_g2p_attr");
    if (rc) return rc;
    GF_REQUIRE(n >= 0 && dx > 0.0f, "gf_k3_g2p_attr: n %d, dx", n);
    if (n == 0) return 0;
    GF_REQUIRE(pos && attr, "gf_k3_g2p_attr: null buffer");
    hipLaunchKernelGGL(k_attr_gather_grids, grid_for(n), dim3(GF_WG), 0, (hipStream_t)stream, pos,
                       attr, channels, n, nx, ny, nz, dx,
                       grids_of(channels, const_cast<float*>(g0), const_cast<float*>(g1),
                                const_cast<float*>(g2), const_cast<float*>(gw)));
    GF_LAUNCH_CHECK("k_attr_gather_grids");
    return 0;
}
