// Pressure projection on the compacted active 8^3 tiles.
//
// Storage: every vector holds 512 entries per active tile (slot-major, cell
// index (li*8+lj)*8+lk), so a wavefront streams one 64-cell li-plane.  Entries
// of non-fluid cells are kept at zero; the 7-point operator
// A x = diag*x - sum_{fluid nb} x of k3_apply_A (solver3d.py:1743-1797) then
// equals diag*x minus the sum over all six neighbours, with diag = number of
// in-grid non-solid neighbours stored in the flags byte.
//
// PCG (solver3d.py:2917-3043 restated): Jacobi preconditioner max(diag,1)
// (k3_compute_diag, 1803-1823), x0 = 0, stop on ||r||_2 < tol*||b||_2,
// alpha = 0 when |pAp| < 1e-30 and beta = 0 when |rz| < 1e-30 (1883-1896).
// Three kernels per iteration; every scalar stays on the device:
//   spmv:   Ap = A p, p.Ap                          -> alpha
//   update: x += alpha p, r -= alpha Ap, z = r/M    -> r.z, r.r, beta, stop flags
//   dir:    p = z + beta p
// Dot products are FP64 with a deterministic two-level reduction.  In the
// mixed-precision mode the vectors are FP32 and an FP64 iterate x64 + x32 is
// kept; the residual is replaced by the FP64 true residual b - A(x64+x32)
// whenever ||r|| has dropped by 10x since the last replacement and before
// convergence is accepted (reliable updates, Sleijpen & van der Vorst 1996).
#include "gf_common.h"
#include <string.h>

namespace {

struct PcgState {
    double rz, pAp, alpha, beta, rr, bb, rr_upd, tol2, delta2, rr_true;
    int done, converged, upd, merge, iters, max_iter, maxed, nupd;
    int k3go, pad0;
    unsigned int counter;
    int pad[5];
};
static_assert(sizeof(PcgState) <= 64 * sizeof(double), "state must fit in scalars");

struct PcgArgs {
    const uint8_t* flags;
    const int32_t* nbr;
    const float* b;
    float* x32; float* r32; float* p32; float* ap32;
    double* x64; double* r64; double* p64; double* ap64;
    float* pc;
    double* partials;
    PcgState* st;
    int nact;
    float rhs_sign;
};

__device__ __forceinline__ int cell_l(int h) {
    const int li = (threadIdx.x >> 6) + 4 * h, lj = (threadIdx.x >> 3) & 7, lk = threadIdx.x & 7;
    return (li * 8 + lj) * 8 + lk;
}

// Sum of the six neighbour values in the order -x,+x,-y,+y,-z,+z of the
// reference stencil; neighbours in inactive tiles read as 0.
template <typename T, typename F>
__device__ __forceinline__ T nb_sum(F get, size_t base, int l, const int* s_nbr) {
    const int li = l >> 6, lj = (l >> 3) & 7, lk = l & 7;
    T s = (T)0;
    s += li > 0 ? get(base + l - 64) : (s_nbr[0] >= 0 ? get((size_t)s_nbr[0] * 512 + l + 448) : (T)0);
    s += li < 7 ? get(base + l + 64) : (s_nbr[1] >= 0 ? get((size_t)s_nbr[1] * 512 + l - 448) : (T)0);
    s += lj > 0 ? get(base + l - 8) : (s_nbr[2] >= 0 ? get((size_t)s_nbr[2] * 512 + l + 56) : (T)0);
    s += lj < 7 ? get(base + l + 8) : (s_nbr[3] >= 0 ? get((size_t)s_nbr[3] * 512 + l - 56) : (T)0);
    s += lk > 0 ? get(base + l - 1) : (s_nbr[4] >= 0 ? get((size_t)s_nbr[4] * 512 + l + 7) : (T)0);
    s += lk < 7 ? get(base + l + 1) : (s_nbr[5] >= 0 ? get((size_t)s_nbr[5] * 512 + l - 7) : (T)0);
    return s;
}

__device__ __forceinline__ void load_nbr(const int32_t* nbr, int slot, int* s_nbr) {
    if (threadIdx.x < 6) s_nbr[threadIdx.x] = nbr[(size_t)slot * 6 + threadIdx.x];
    __syncthreads();
}

template <typename T> __device__ __forceinline__ T* vx(const PcgArgs& a);
template <> __device__ __forceinline__ float* vx<float>(const PcgArgs& a) { return a.x32; }
template <> __device__ __forceinline__ double* vx<double>(const PcgArgs& a) { return a.x64; }
template <typename T> __device__ __forceinline__ T* vr(const PcgArgs& a);
template <> __device__ __forceinline__ float* vr<float>(const PcgArgs& a) { return a.r32; }
template <> __device__ __forceinline__ double* vr<double>(const PcgArgs& a) { return a.r64; }
template <typename T> __device__ __forceinline__ T* vp(const PcgArgs& a);
template <> __device__ __forceinline__ float* vp<float>(const PcgArgs& a) { return a.p32; }
template <> __device__ __forceinline__ double* vp<double>(const PcgArgs& a) { return a.p64; }
template <typename T> __device__ __forceinline__ T* vap(const PcgArgs& a);
template <> __device__ __forceinline__ float* vap<float>(const PcgArgs& a) { return a.ap32; }
template <> __device__ __forceinline__ double* vap<double>(const PcgArgs& a) { return a.ap64; }

// x = 0, r = rhs_sign*b, p = z = r/M, rz = r.z, bb = r.r
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_setup(PcgArgs a, double tol2, double delta2,
                                                     int max_iter) {
    const int slot = gf::xcd_remap(blockIdx.x, a.nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
    T* r = vr<T>(a);
    T* p = vp<T>(a);
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const size_t idx = base + cell_l(h);
        const uint8_t f = a.flags[idx];
        T rv = (T)0, z = (T)0;
        if (f & 8) {
            rv = (T)(a.rhs_sign * a.b[idx]);
            const int d = f & 7;
            z = rv / (T)(d > 1 ? d : 1);
        }
        r[idx] = rv;
        p[idx] = z;
        a.x64[idx] = 0.0;
        if (MIXED) a.x32[idx] = 0.0f;
        v[0] += (double)rv * (double)z;
        v[1] += (double)rv * (double)rv;
    }
    double tot[2];
    if (gf::wg_reduce_last<2>(v, a.partials, gridDim.x, &a.st->counter, tot) && threadIdx.x == 0) {
        PcgState* st = a.st;
        st->rz = tot[0];
        st->bb = tot[1];
        st->rr = tot[1];
        st->rr_upd = tot[1];
        st->rr_true = tot[1];
        st->tol2 = tol2;
        st->delta2 = delta2;
        st->max_iter = max_iter;
        st->iters = 0;
        st->maxed = 0;
        st->upd = 0;
        st->merge = 0;
        st->nupd = 0;
        st->k3go = 0;
        const int zero = tot[1] < 1e-30;
        st->done = zero || max_iter <= 0;
        st->converged = zero;
    }
}

// Ap = A p, pAp -> alpha
template <typename T>
__global__ __launch_bounds__(GF_WG) void k_pcg_spmv(PcgArgs a) {
    __shared__ int s_nbr[6];
    if (a.st->done) return;
    const int slot = gf::xcd_remap(blockIdx.x, a.nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
    load_nbr(a.nbr, slot, s_nbr);
    const T* __restrict__ p = vp<T>(a);
    T* __restrict__ ap = vap<T>(a);
    auto get = [&](size_t i) -> T { return p[i]; };
    double v[1] = {0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int l = cell_l(h);
        const size_t idx = base + l;
        const uint8_t f = a.flags[idx];
        if (f & 8) {
            const T xc = p[idx];
            const T s = nb_sum<T>(get, base, l, s_nbr);
            const T y = (T)(f & 7) * xc - s;
            ap[idx] = y;
            v[0] += (double)xc * (double)y;
        }
    }
    double tot[1];
    if (gf::wg_reduce_last<1>(v, a.partials, gridDim.x, &a.st->counter, tot) && threadIdx.x == 0) {
        PcgState* st = a.st;
        st->pAp = tot[0];
        st->alpha = fabs(tot[0]) < 1e-30 ? 0.0 : st->rz / tot[0];
    }
}

// x += alpha p, r -= alpha Ap; r.z, r.r -> beta, stop / replacement flags
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_update(PcgArgs a) {
    PcgState* st = a.st;
    if (st->done) return;
    const int slot = gf::xcd_remap(blockIdx.x, a.nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
    const double alpha = st->alpha;
    const T af = (T)alpha;
    const int merge = MIXED ? st->merge : 0;
    T* __restrict__ r = vr<T>(a);
    const T* __restrict__ p = vp<T>(a);
    const T* __restrict__ ap = vap<T>(a);
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const size_t idx = base + cell_l(h);
        const uint8_t f = a.flags[idx];
        if (f & 8) {
            const T pv = p[idx];
            if (MIXED) {
                float xo = a.x32[idx];
                if (merge) {
                    a.x64[idx] += (double)xo;
                    xo = 0.0f;
                }
                a.x32[idx] = xo + (float)af * (float)pv;
            } else {
                a.x64[idx] += (double)af * (double)pv;
            }
            const T rv = r[idx] - af * ap[idx];
            r[idx] = rv;
            const int d = f & 7;
            const T z = rv / (T)(d > 1 ? d : 1);
            v[0] += (double)rv * (double)z;
            v[1] += (double)rv * (double)rv;
        }
    }
    double tot[2];
    if (gf::wg_reduce_last<2>(v, a.partials, gridDim.x, &st->counter, tot) && threadIdx.x == 0) {
        const double rz_old = st->rz;
        st->beta = fabs(rz_old) < 1e-30 ? 0.0 : tot[0] / rz_old;
        st->rz = tot[0];
        st->rr = tot[1];
        st->iters += 1;
        st->merge = 0;
        st->k3go = 1;
        if (tot[1] < st->tol2 * st->bb) {
            st->done = 1;
            if (!MIXED) { st->converged = 1; st->rr_true = tot[1]; }
        } else if (st->iters >= st->max_iter) {
            st->done = 1;
            st->maxed = 1;
            if (!MIXED) st->rr_true = tot[1];
        } else if (MIXED && tot[1] < st->delta2 * st->rr_upd) {
            st->upd = 1;
        }
    }
}

// p = z + beta p (runs once after each executed update kernel)
template <typename T>
__global__ __launch_bounds__(GF_WG) void k_pcg_dir(PcgArgs a) {
    __shared__ int s_last;
    PcgState* st = a.st;
    if (!st->k3go) return;
    const int slot = gf::xcd_remap(blockIdx.x, a.nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
    const T bf = (T)st->beta;
    const T* __restrict__ r = vr<T>(a);
    T* __restrict__ p = vp<T>(a);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const size_t idx = base + cell_l(h);
        const uint8_t f = a.flags[idx];
        if (f & 8) {
            const int d = f & 7;
            const T z = r[idx] / (T)(d > 1 ? d : 1);
            p[idx] = z + bf * p[idx];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned int ticket = atomicAdd(&st->counter, 1u);
        s_last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        st->k3go = 0;
        st->counter = 0u;
    }
}

// Residual replacement: r = rhs - A (x64 + x32) in FP64 (mixed mode only).
__global__ __launch_bounds__(GF_WG) void k_pcg_replace(PcgArgs a) {
    __shared__ int s_nbr[6];
    PcgState* st = a.st;
    if (!(st->upd || (st->done && !st->converged))) return;
    const int slot = gf::xcd_remap(blockIdx.x, a.nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
    load_nbr(a.nbr, slot, s_nbr);
    const double* __restrict__ x64 = a.x64;
    const float* __restrict__ x32 = a.x32;
    auto get = [&](size_t i) -> double { return x64[i] + (double)x32[i]; };
    double v[2] = {0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int l = cell_l(h);
        const size_t idx = base + l;
        const uint8_t f = a.flags[idx];
        if (f & 8) {
            const int d = f & 7;
            const double ax = (double)d * get(idx) - nb_sum<double>(get, base, l, s_nbr);
            const double rv = (double)(a.rhs_sign * a.b[idx]) - ax;
            const float r32 = (float)rv;
            a.r32[idx] = r32;
            const float z = r32 / (float)(d > 1 ? d : 1);
            v[0] += rv * rv;
            v[1] += (double)r32 * (double)z;
        }
    }
    double tot[2];
    if (gf::wg_reduce_last<2>(v, a.partials, gridDim.x, &st->counter, tot) && threadIdx.x == 0) {
        st->rr_true = tot[0];
        st->rr = tot[0];
        st->rr_upd = tot[0];
        st->rz = tot[1];
        st->upd = 0;
        st->merge = 1;
        st->nupd += 1;
        if (tot[0] < st->tol2 * st->bb) {
            st->converged = 1;
            st->done = 1;
        } else if (st->iters >= st->max_iter) {
            st->maxed = 1;
            st->done = 1;
        } else {
            st->done = 0;
        }
    }
}

// pc = FP32 result (x64 [+ x32]); non-fluid cells 0.
template <bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_final(PcgArgs a) {
    const int slot = gf::xcd_remap(blockIdx.x, a.nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const size_t idx = base + cell_l(h);
        const uint8_t f = a.flags[idx];
        float out = 0.0f;
        if (f & 8) out = MIXED ? (float)(a.x64[idx] + (double)a.x32[idx]) : (float)a.x64[idx];
        a.pc[idx] = out;
    }
}

// Jacobi sweep (solver3d.py:878-938) on compact tiles.
__global__ __launch_bounds__(GF_WG) void k_jacobi(const uint8_t* __restrict__ flags,
                                                  const int32_t* __restrict__ nbr,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ pin,
                                                  float* __restrict__ pout, int nact) {
    __shared__ int s_nbr[6];
    const int slot = gf::xcd_remap(blockIdx.x, nact);
    const size_t base = (size_t)slot * GF_TILE_CELLS;
    load_nbr(nbr, slot, s_nbr);
    auto get = [&](size_t i) -> float { return pin[i]; };
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int l = cell_l(h);
        const size_t idx = base + l;
        const uint8_t f = flags[idx];
        if (f & 8) {
            const int d = f & 7;
            float out = 0.0f;
            if (d > 0) {
                const float s = nb_sum<float>(get, base, l, s_nbr);
                out = (s - b[idx]) / (float)d;
            }
            pout[idx] = out;
        }
    }
}

__global__ void k_fill_f32(float* a, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

namespace gf {
int scene_check(const gf_scene_t* s, const char* who);
}

extern "C" int gf_scene_jacobi(const gf_scene_t* s, int n_active, int iters, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_jacobi");
    if (rc) return rc;
    const int ntiles = s->tnx * s->tny * s->tnz;
    GF_REQUIRE(n_active >= 0 && n_active <= ntiles, "gf_scene_jacobi: n_active %d", n_active);
    GF_REQUIRE(iters >= 0, "gf_scene_jacobi: iters %d", iters);
    GF_REQUIRE(s->flags && s->nbr && s->b && s->pc && s->ptmp, "gf_scene_jacobi: buffers not set");
    if (n_active == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)n_active * GF_TILE_CELLS;
    const int fb = (int)((n + GF_WG - 1) / GF_WG);
    hipLaunchKernelGGL(k_fill_f32, dim3(fb), dim3(GF_WG), 0, st, s->pc, n, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(fb), dim3(GF_WG), 0, st, s->ptmp, n, 0.0f);
    for (int k = 0; k < iters; ++k) {
        const bool to_pc = ((iters - 1 - k) % 2) == 0;
        float* out = to_pc ? s->pc : s->ptmp;
        const float* in = to_pc ? s->ptmp : s->pc;
        hipLaunchKernelGGL(k_jacobi, dim3(n_active), dim3(GF_WG), 0, st, s->flags, s->nbr, s->b,
                           in, out, n_active);
    }
    GF_LAUNCH_CHECK("k_jacobi");
    return 0;
}

extern "C" int gf_scene_pcg_sync(const gf_scene_t* s, int n_active, int precision, double tol,
                                 int max_iter, int check_every, int rhs_sign, int* iters,
                                 double* rel_residual, int* replacements, double* kernel_ms,
                                 void* stream) {
    int rc = gf::scene_check(s, "gf_scene_pcg_sync");
    if (rc) return rc;
    const int ntiles = s->tnx * s->tny * s->tnz;
    GF_REQUIRE(n_active >= 0 && n_active <= ntiles, "gf_scene_pcg_sync: n_active %d", n_active);
    GF_REQUIRE(precision == GF_PCG_MIXED || precision == GF_PCG_FP64,
               "gf_scene_pcg_sync: precision %d", precision);
    GF_REQUIRE(rhs_sign == 1 || rhs_sign == -1, "gf_scene_pcg_sync: rhs_sign %d", rhs_sign);
    GF_REQUIRE(tol >= 0.0 && max_iter >= 0, "gf_scene_pcg_sync: tol/max_iter");
    GF_REQUIRE(s->flags && s->nbr && s->b && s->pc && s->partials && s->scalars && s->x64,
               "gf_scene_pcg_sync: buffers not set");
    if (precision == GF_PCG_MIXED)
        GF_REQUIRE(s->x32 && s->r32 && s->p32 && s->ap32, "gf_scene_pcg_sync: FP32 vectors not set");
    else
        GF_REQUIRE(s->r64 && s->p64 && s->ap64, "gf_scene_pcg_sync: FP64 vectors not set");
    if (check_every < 1) check_every = 1;
    if (iters) *iters = 0;
    if (rel_residual) *rel_residual = 0.0;
    if (replacements) *replacements = 0;
    hipStream_t st = (hipStream_t)stream;
    if (n_active == 0) return 0;

    PcgArgs a;
    a.flags = s->flags; a.nbr = s->nbr; a.b = s->b;
    a.x32 = s->x32; a.r32 = s->r32; a.p32 = s->p32; a.ap32 = s->ap32;
    a.x64 = s->x64; a.r64 = s->r64; a.p64 = s->p64; a.ap64 = s->ap64;
    a.pc = s->pc; a.partials = s->partials; a.st = (PcgState*)s->scalars;
    a.nact = n_active; a.rhs_sign = (float)rhs_sign;
    const dim3 grid(n_active), block(GF_WG);
    const bool mixed = precision == GF_PCG_MIXED;
    const double tol2 = tol * tol, delta2 = 0.01;

    // The reduction counter must start at zero; clear it once per solve.
    hipError_t e = hipMemsetAsync(&a.st->counter, 0, sizeof(unsigned int), st);
    if (e != hipSuccess)
        return gf::set_error((int)e, "gf_scene_pcg_sync: counter reset: %s", hipGetErrorString(e));
    if (mixed)
        hipLaunchKernelGGL((k_pcg_setup<float, true>), grid, block, 0, st, a, tol2, delta2, max_iter);
    else
        hipLaunchKernelGGL((k_pcg_setup<double, false>), grid, block, 0, st, a, tol2, delta2, max_iter);
    GF_LAUNCH_CHECK("k_pcg_setup");

    // Optional per-kernel timing of the iterations that did work: events
    // around every launch of a chunk, read back at the chunk's host sync.
    const bool timed = kernel_ms != nullptr;
    hipEvent_t ev[4 * 64 + 2];
    int nev = 0;
    if (timed) {
        if (check_every > 64) check_every = 64;
        for (int i = 0; i < 4; ++i) kernel_ms[i] = 0.0;
        nev = 4 * check_every + 2;
        for (int i = 0; i < nev; ++i) {
            e = hipEventCreate(&ev[i]);
            if (e != hipSuccess) {
                for (int j = 0; j < i; ++j) hipEventDestroy(ev[j]);
                return gf::set_error((int)e, "gf_scene_pcg_sync: event: %s", hipGetErrorString(e));
            }
        }
    }
    PcgState hs;
    memset(&hs, 0, sizeof(hs));
    int launched = 0, iters_before = 0, nupd_before = 0;
    rc = 0;
    for (;;) {
        const int chunk = check_every;
        for (int j = 0; j < chunk; ++j) {
            if (timed) hipEventRecord(ev[4 * j], st);
            if (mixed) hipLaunchKernelGGL(k_pcg_spmv<float>, grid, block, 0, st, a);
            else hipLaunchKernelGGL(k_pcg_spmv<double>, grid, block, 0, st, a);
            if (timed) hipEventRecord(ev[4 * j + 1], st);
            if (mixed) hipLaunchKernelGGL((k_pcg_update<float, true>), grid, block, 0, st, a);
            else hipLaunchKernelGGL((k_pcg_update<double, false>), grid, block, 0, st, a);
            if (timed) hipEventRecord(ev[4 * j + 2], st);
            if (mixed) hipLaunchKernelGGL(k_pcg_dir<float>, grid, block, 0, st, a);
            else hipLaunchKernelGGL(k_pcg_dir<double>, grid, block, 0, st, a);
            if (timed) hipEventRecord(ev[4 * j + 3], st);
        }
        launched += chunk;
        if (mixed) {
            if (timed) hipEventRecord(ev[4 * chunk], st);
            hipLaunchKernelGGL(k_pcg_replace, grid, block, 0, st, a);
            if (timed) hipEventRecord(ev[4 * chunk + 1], st);
        }
        rc = gf::check_launch("pcg iteration");
        if (rc) break;
        e = hipMemcpyAsync(&hs, a.st, sizeof(PcgState), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = gf::set_error((int)e, "gf_scene_pcg_sync: readback: %s", hipGetErrorString(e));
            break;
        }
        if (timed) {
            const int real = hs.iters - iters_before;
            for (int j = 0; j < real && j < chunk; ++j) {
                float t0 = 0.f, t1 = 0.f, t2 = 0.f;
                hipEventElapsedTime(&t0, ev[4 * j], ev[4 * j + 1]);
                hipEventElapsedTime(&t1, ev[4 * j + 1], ev[4 * j + 2]);
                hipEventElapsedTime(&t2, ev[4 * j + 2], ev[4 * j + 3]);
                kernel_ms[0] += t0;
                kernel_ms[1] += t1;
                kernel_ms[2] += t2;
            }
            if (mixed && hs.nupd > nupd_before) {
                float t3 = 0.f;
                hipEventElapsedTime(&t3, ev[4 * chunk], ev[4 * chunk + 1]);
                kernel_ms[3] += t3;
            }
        }
        iters_before = hs.iters;
        nupd_before = hs.nupd;
        if (hs.done) break;
        if (hs.iters >= max_iter || launched > max_iter + 2 * chunk) break;
    }
    for (int i = 0; i < nev; ++i) hipEventDestroy(ev[i]);
    if (rc) return rc;
    if (mixed)
        hipLaunchKernelGGL(k_pcg_final<true>, grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(k_pcg_final<false>, grid, block, 0, st, a);
    GF_LAUNCH_CHECK("k_pcg_final");
    if (iters) *iters = hs.iters;
    if (rel_residual) *rel_residual = hs.bb > 0.0 ? sqrt(hs.rr_true / hs.bb) : 0.0;
    if (replacements) *replacements = hs.nupd;
    return 0;
}
