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

// Solver scalars.  Kernels form a fixed launch sequence numbered L; launch L
// reads entry (L-1)&1 and writes entry L&1 (workgroup 0 only), so no kernel
// reads a field that it also writes.
struct PcgLog {
    double rz, rr, bb, alpha, beta, rr_upd, rr_true, pap;
    int iters, done, converged, maxed, upd, merge, nupd, pending;
    double pad[4];
};
static_assert(sizeof(PcgLog) == 16 * sizeof(double), "log entry is 16 doubles");

// Per-workgroup partial sums, PCG_GMAX entries per kind.
enum { P_PAP = 0, P_RZ, P_RR, P_BB, P_RRT, P_RZT, P_KINDS };
constexpr int PCG_GMAX = 2048;

struct PcgArgs {
    const uint8_t* flags;
    const int32_t* nbr;
    const float* b;
    float* x32; float* r32; float* p32; float* ap32;
    double* x64; double* r64; double* p64; double* ap64;
    float* pc;
    double* part;        // [P_KINDS][PCG_GMAX]
    PcgLog* log;         // [2]
    int nact;            // active tiles
    int G;               // workgroups; workgroup b owns tiles [b*nact/G, (b+1)*nact/G)
    float rhs_sign;
    double tol2, delta2;
    int max_iter;
};

// ---- quad addressing ----------------------------------------------------
// A workgroup processes two tiles per pass; thread -> (tile half, quad q),
// quad q covers cells (li, lj, 4*kh .. 4*kh+3) of its tile, i.e. one aligned
// 16-byte (FP32) / 32-byte (FP64) vector.
struct Quad {
    int li, lj, kh, l0;
};

__device__ __forceinline__ Quad quad_of() {
    const int q = threadIdx.x & 127;
    Quad Q;
    Q.li = q >> 4;
    Q.lj = (q >> 1) & 7;
    Q.kh = q & 1;
    Q.l0 = (Q.li * 8 + Q.lj) * 8 + Q.kh * 4;
    return Q;
}

template <typename T> struct V4;
template <> struct V4<float> { using t = float4; };
template <> struct V4<double> { using t = double4; };

template <typename T>
__device__ __forceinline__ void ld4(const T* p, T (&o)[4]) {
    const typename V4<T>::t v = *reinterpret_cast<const typename V4<T>::t*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}

template <typename T>
__device__ __forceinline__ void st4(T* p, const T (&o)[4]) {
    typename V4<T>::t v;
    v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
    *reinterpret_cast<typename V4<T>::t*>(p) = v;
}

__device__ __forceinline__ uint32_t ld_flags(const uint8_t* f, size_t idx) {
    return *reinterpret_cast<const uint32_t*>(f + idx);
}
__device__ __forceinline__ int fl_fluid(uint32_t fw, int e) { return (fw >> (8 * e + 3)) & 1; }
__device__ __forceinline__ int fl_diag(uint32_t fw, int e) { return (fw >> (8 * e)) & 7; }

// Neighbour values of a quad: the four -x/+x/-y/+y vectors and the two z
// values that lie outside the quad (-z of its first cell, +z of its last).
// Neighbours in inactive tiles read as 0; every load uses a valid address.
template <typename T>
struct Nb4 {
    T xm[4], xp[4], ym[4], yp[4];
    T zm0, zp3;
};

template <typename T>
__device__ __forceinline__ void load_nb4(const T* __restrict__ v, size_t base, const Quad& Q,
                                         const int32_t* nb, Nb4<T>& o) {
    const size_t own = base + Q.l0;
    bool ok;
    size_t a;
    ok = Q.li > 0 || nb[0] >= 0;
    a = Q.li > 0 ? own - 64 : (nb[0] >= 0 ? (size_t)nb[0] * 512 + Q.l0 + 448 : own);
    ld4(v + a, o.xm);
    if (!ok) for (int e = 0; e < 4; ++e) o.xm[e] = (T)0;
    ok = Q.li < 7 || nb[1] >= 0;
    a = Q.li < 7 ? own + 64 : (nb[1] >= 0 ? (size_t)nb[1] * 512 + Q.l0 - 448 : own);
    ld4(v + a, o.xp);
    if (!ok) for (int e = 0; e < 4; ++e) o.xp[e] = (T)0;
    ok = Q.lj > 0 || nb[2] >= 0;
    a = Q.lj > 0 ? own - 8 : (nb[2] >= 0 ? (size_t)nb[2] * 512 + Q.l0 + 56 : own);
    ld4(v + a, o.ym);
    if (!ok) for (int e = 0; e < 4; ++e) o.ym[e] = (T)0;
    ok = Q.lj < 7 || nb[3] >= 0;
    a = Q.lj < 7 ? own + 8 : (nb[3] >= 0 ? (size_t)nb[3] * 512 + Q.l0 - 56 : own);
    ld4(v + a, o.yp);
    if (!ok) for (int e = 0; e < 4; ++e) o.yp[e] = (T)0;
    ok = Q.kh == 1 || nb[4] >= 0;
    a = Q.kh == 1 ? own - 1 : (nb[4] >= 0 ? (size_t)nb[4] * 512 + Q.l0 + 7 : own);
    o.zm0 = ok ? v[a] : (T)0;
    ok = Q.kh == 0 || nb[5] >= 0;
    a = Q.kh == 0 ? own + 4 : (nb[5] >= 0 ? (size_t)nb[5] * 512 + Q.l0 - 4 : own);
    o.zp3 = ok ? v[a] : (T)0;
}

// Neighbour sum of quad element e in the reference order -x,+x,-y,+y,-z,+z.
template <typename T>
__device__ __forceinline__ T nb_sum4(const Nb4<T>& o, const T (&c)[4], int e) {
    T s = o.xm[e];
    s += o.xp[e];
    s += o.ym[e];
    s += o.yp[e];
    s += e == 0 ? o.zm0 : c[e - 1];
    s += e == 3 ? o.zp3 : c[e + 1];
    return s;
}

// Tile range of this workgroup; consecutive logical workgroups sit on the same
// XCD so that neighbouring tiles share an L2.
__device__ __forceinline__ void tile_range(const PcgArgs& a, int& t0, int& t1, int& b) {
    b = gf::xcd_remap(blockIdx.x, a.G);
    t0 = (int)(((long long)b * a.nact) / a.G);
    t1 = (int)(((long long)(b + 1) * a.nact) / a.G);
}

// Workgroup sum of one value per thread, in a fixed order; result in all threads.
__device__ __forceinline__ double wg_sum(double v, double* s_red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    v = gf::wave_sum(v);
    __syncthreads();
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < GF_WG / GF_WAVE; ++i) t += s_red[i];
    return t;
}

// Sum of the G partials of one kind (deterministic, identical in every
// workgroup).  Written by the previous kernel, so visible at this launch.
__device__ __forceinline__ double sum_partials(const double* part, int kind, int G,
                                               double* s_red) {
    double v = 0.0;
    for (int i = threadIdx.x; i < G; i += GF_WG) v += part[kind * PCG_GMAX + i];
    return wg_sum(v, s_red);
}

__device__ __forceinline__ void put_partial(double* part, int kind, int b, double v,
                                            double* s_red) {
    const double t = wg_sum(v, s_red);
    if (threadIdx.x == 0) part[kind * PCG_GMAX + b] = t;
}

template <typename T> __device__ __forceinline__ T* vr(const PcgArgs& a);
template <> __device__ __forceinline__ float* vr<float>(const PcgArgs& a) { return a.r32; }
template <> __device__ __forceinline__ double* vr<double>(const PcgArgs& a) { return a.r64; }
template <typename T> __device__ __forceinline__ T* vp(const PcgArgs& a);
template <> __device__ __forceinline__ float* vp<float>(const PcgArgs& a) { return a.p32; }
template <> __device__ __forceinline__ double* vp<double>(const PcgArgs& a) { return a.p64; }
template <typename T> __device__ __forceinline__ T* vap(const PcgArgs& a);
template <> __device__ __forceinline__ float* vap<float>(const PcgArgs& a) { return a.ap32; }
template <> __device__ __forceinline__ double* vap<double>(const PcgArgs& a) { return a.ap64; }

enum { PEND_NONE = 0, PEND_SETUP = 1, PEND_REPLACE = 2 };

#define FOR_QUADS(a)                                                        \
    int t0_, t1_, b;                                                        \
    tile_range(a, t0_, t1_, b);                                             \
    const Quad Q = quad_of();                                               \
    for (int tb_ = t0_; tb_ < t1_; tb_ += 2)                                \
        if (const int t = tb_ + (threadIdx.x >> 7); t < t1_)

// L = 0: x = 0, r = rhs_sign*b, p = z = r/M; partials r.z and r.r.
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_setup(PcgArgs a) {
    __shared__ double s_red[GF_WG / GF_WAVE];
    T* r = vr<T>(a);
    T* p = vp<T>(a);
    double rz = 0.0, rr = 0.0;
    FOR_QUADS(a) {
        const size_t idx = (size_t)t * GF_TILE_CELLS + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        float bv[4];
        ld4(a.b + idx, bv);
        T rv[4], z[4], zero[4] = {(T)0, (T)0, (T)0, (T)0};
        double zd[4] = {0.0, 0.0, 0.0, 0.0};
        for (int e = 0; e < 4; ++e) {
            rv[e] = fl_fluid(fw, e) ? (T)(a.rhs_sign * bv[e]) : (T)0;
            const int d = fl_diag(fw, e);
            z[e] = rv[e] / (T)(d > 1 ? d : 1);
            rz += (double)rv[e] * (double)z[e];
            rr += (double)rv[e] * (double)rv[e];
        }
        st4(r + idx, rv);
        st4(p + idx, z);
        st4(a.x64 + idx, zd);
        if (MIXED) {
            float zf[4] = {0.f, 0.f, 0.f, 0.f};
            st4(a.x32 + idx, zf);
        }
        (void)zero;
    }
    put_partial(a.part, P_RZ, b, rz, s_red);
    put_partial(a.part, P_BB, b, rr, s_red);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PcgLog e = {};
        e.pending = PEND_SETUP;
        a.log[0] = e;
    }
}

// One-workgroup finalisation after the setup or a residual replacement.
__global__ __launch_bounds__(GF_WG) void k_pcg_fin(PcgArgs a, int L) {
    __shared__ double s_red[GF_WG / GF_WAVE];
    PcgLog e = a.log[(L - 1) & 1];
    if (e.pending == PEND_SETUP) {
        const double rz = sum_partials(a.part, P_RZ, a.G, s_red);
        const double bb = sum_partials(a.part, P_BB, a.G, s_red);
        e.rz = rz; e.bb = bb; e.rr = bb; e.rr_upd = bb; e.rr_true = bb;
        e.iters = 0; e.maxed = 0; e.upd = 0; e.merge = 0; e.nupd = 0;
        const int zero = bb < 1e-30;
        e.converged = zero;
        e.done = zero || a.max_iter <= 0;
    } else if (e.pending == PEND_REPLACE) {
        const double rrt = sum_partials(a.part, P_RRT, a.G, s_red);
        const double rz = sum_partials(a.part, P_RZT, a.G, s_red);
        e.rr_true = rrt; e.rr = rrt; e.rr_upd = rrt; e.rz = rz;
        e.upd = 0; e.merge = 1; e.nupd += 1;
        if (rrt < a.tol2 * e.bb) { e.converged = 1; e.done = 1; }
        else if (e.iters >= a.max_iter) { e.maxed = 1; e.done = 1; }
        else e.done = 0;
    }
    e.pending = PEND_NONE;
    if (threadIdx.x == 0) a.log[L & 1] = e;
}

__device__ __forceinline__ bool pass_through(const PcgArgs& a, const PcgLog& e, int L) {
    if (!e.done) return false;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.log[L & 1] = e;
    return true;
}

// Ap = A p; partial p.Ap.
template <typename T>
__global__ __launch_bounds__(GF_WG) void k_pcg_spmv(PcgArgs a, int L) {
    __shared__ double s_red[GF_WG / GF_WAVE];
    const PcgLog e = a.log[(L - 1) & 1];
    if (pass_through(a, e, L)) return;
    const T* __restrict__ p = vp<T>(a);
    T* __restrict__ ap = vap<T>(a);
    double pap = 0.0;
    FOR_QUADS(a) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        T c[4], y[4];
        ld4(p + idx, c);
        Nb4<T> o;
        load_nb4(p, base, Q, a.nbr + (size_t)t * 6, o);
        for (int k = 0; k < 4; ++k) {
            const T s = nb_sum4(o, c, k);
            y[k] = fl_fluid(fw, k) ? (T)fl_diag(fw, k) * c[k] - s : (T)0;
            pap += (double)c[k] * (double)y[k];
        }
        st4(ap + idx, y);
    }
    put_partial(a.part, P_PAP, b, pap, s_red);
    if (blockIdx.x == 0 && threadIdx.x == 0) a.log[L & 1] = e;
}

// alpha = rz / pAp; x += alpha p, r -= alpha Ap; partials r.z, r.r.
// Non-fluid entries of p, Ap, r and x are 0 and stay 0.
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_update(PcgArgs a, int L) {
    __shared__ double s_red[GF_WG / GF_WAVE];
    PcgLog e = a.log[(L - 1) & 1];
    if (pass_through(a, e, L)) return;
    const double pap_t = sum_partials(a.part, P_PAP, a.G, s_red);
    const double alpha = fabs(pap_t) < 1e-30 ? 0.0 : e.rz / pap_t;
    const T af = (T)alpha;
    const int merge = MIXED ? e.merge : 0;
    T* __restrict__ r = vr<T>(a);
    const T* __restrict__ p = vp<T>(a);
    const T* __restrict__ ap = vap<T>(a);
    double rz = 0.0, rr = 0.0;
    FOR_QUADS(a) {
        const size_t idx = (size_t)t * GF_TILE_CELLS + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        T pv[4], apv[4], rv[4];
        ld4(p + idx, pv);
        ld4(ap + idx, apv);
        ld4(r + idx, rv);
        if (MIXED) {
            float xo[4];
            ld4(a.x32 + idx, xo);
            if (merge) {
                double xd[4];
                ld4(a.x64 + idx, xd);
                for (int k = 0; k < 4; ++k) {
                    xd[k] += (double)xo[k];
                    xo[k] = 0.0f;
                }
                st4(a.x64 + idx, xd);
            }
            for (int k = 0; k < 4; ++k) xo[k] = xo[k] + (float)af * (float)pv[k];
            st4(a.x32 + idx, xo);
        } else {
            double xd[4];
            ld4(a.x64 + idx, xd);
            for (int k = 0; k < 4; ++k) xd[k] += (double)af * (double)pv[k];
            st4(a.x64 + idx, xd);
        }
        for (int k = 0; k < 4; ++k) {
            rv[k] = rv[k] - af * apv[k];
            const int d = fl_diag(fw, k);
            const T z = rv[k] / (T)(d > 1 ? d : 1);
            rz += (double)rv[k] * (double)z;
            rr += (double)rv[k] * (double)rv[k];
        }
        st4(r + idx, rv);
    }
    put_partial(a.part, P_RZ, b, rz, s_red);
    put_partial(a.part, P_RR, b, rr, s_red);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        e.alpha = alpha;
        e.pap = pap_t;
        e.merge = 0;
        a.log[L & 1] = e;
    }
}

// beta = rz_new / rz; p = z + beta p; convergence / replacement decisions.
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_dir(PcgArgs a, int L) {
    __shared__ double s_red[GF_WG / GF_WAVE];
    PcgLog e = a.log[(L - 1) & 1];
    if (pass_through(a, e, L)) return;
    const double rz_new = sum_partials(a.part, P_RZ, a.G, s_red);
    const double rr = sum_partials(a.part, P_RR, a.G, s_red);
    const double beta = fabs(e.rz) < 1e-30 ? 0.0 : rz_new / e.rz;
    const T bf = (T)beta;
    const T* __restrict__ r = vr<T>(a);
    T* __restrict__ p = vp<T>(a);
    FOR_QUADS(a) {
        const size_t idx = (size_t)t * GF_TILE_CELLS + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        T rv[4], pv[4];
        ld4(r + idx, rv);
        ld4(p + idx, pv);
        for (int k = 0; k < 4; ++k) {
            const int d = fl_diag(fw, k);
            const T z = rv[k] / (T)(d > 1 ? d : 1);
            pv[k] = z + bf * pv[k];
        }
        st4(p + idx, pv);
    }
    (void)b;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        e.beta = beta;
        e.rz = rz_new;
        e.rr = rr;
        e.iters += 1;
        if (rr < a.tol2 * e.bb) {
            e.done = 1;
            if (!MIXED) { e.converged = 1; e.rr_true = rr; }
        } else if (e.iters >= a.max_iter) {
            e.done = 1;
            e.maxed = 1;
            if (!MIXED) e.rr_true = rr;
        } else if (MIXED && rr < a.delta2 * e.rr_upd) {
            e.upd = 1;
        }
        a.log[L & 1] = e;
    }
}

// Residual replacement (mixed mode): r = rhs - A (x64 + x32) in FP64 when a
// replacement is due or convergence is to be confirmed; partials of the true
// r.r and of r.z.  Finalised by k_pcg_fin.
__global__ __launch_bounds__(GF_WG) void k_pcg_replace(PcgArgs a, int L) {
    __shared__ double s_red[GF_WG / GF_WAVE];
    PcgLog e = a.log[(L - 1) & 1];
    const bool run = e.upd || (e.done && !e.converged);
    if (!run) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.log[L & 1] = e;
        return;
    }
    double rrt = 0.0, rz = 0.0;
    FOR_QUADS(a) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        const int32_t* nb = a.nbr + (size_t)t * 6;
        double cd[4];
        float cf[4], bv[4];
        ld4(a.x64 + idx, cd);
        ld4(a.x32 + idx, cf);
        ld4(a.b + idx, bv);
        Nb4<double> od;
        Nb4<float> of;
        load_nb4(a.x64, base, Q, nb, od);
        load_nb4(a.x32, base, Q, nb, of);
        double c[4];
        for (int k = 0; k < 4; ++k) {
            c[k] = cd[k] + (double)cf[k];
            od.xm[k] += (double)of.xm[k];
            od.xp[k] += (double)of.xp[k];
            od.ym[k] += (double)of.ym[k];
            od.yp[k] += (double)of.yp[k];
        }
        od.zm0 += (double)of.zm0;
        od.zp3 += (double)of.zp3;
        float r32[4];
        for (int k = 0; k < 4; ++k) {
            float rk = 0.0f;
            if (fl_fluid(fw, k)) {
                const int d = fl_diag(fw, k);
                const double ax = (double)d * c[k] - nb_sum4(od, c, k);
                const double rv = (double)(a.rhs_sign * bv[k]) - ax;
                rk = (float)rv;
                const float z = rk / (float)(d > 1 ? d : 1);
                rrt += rv * rv;
                rz += (double)rk * (double)z;
            }
            r32[k] = rk;
        }
        st4(a.r32 + idx, r32);
    }
    put_partial(a.part, P_RRT, b, rrt, s_red);
    put_partial(a.part, P_RZT, b, rz, s_red);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        e.pending = PEND_REPLACE;
        a.log[L & 1] = e;
    }
}

// pc = FP32 result (x64 [+ x32]); non-fluid cells hold 0 in both.
template <bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_final(PcgArgs a) {
    FOR_QUADS(a) {
        const size_t idx = (size_t)t * GF_TILE_CELLS + Q.l0;
        double xd[4];
        ld4(a.x64 + idx, xd);
        float out[4];
        if (MIXED) {
            float xf[4];
            ld4(a.x32 + idx, xf);
            for (int k = 0; k < 4; ++k) out[k] = (float)(xd[k] + (double)xf[k]);
        } else {
            for (int k = 0; k < 4; ++k) out[k] = (float)xd[k];
        }
        st4(a.pc + idx, out);
    }
    (void)b;
}

// Jacobi sweep (solver3d.py:878-938) on compact tiles: two tiles per
// workgroup, p_out = (sum of fluid-neighbour p_in - div) / diag on fluid cells
// with diag > 0, else 0 (non-fluid p_in entries are 0, so the neighbour sum in
// the reference order equals the reference's fluid-only sum).
__global__ __launch_bounds__(GF_WG) void k_jacobi(const uint8_t* __restrict__ flags,
                                                  const int32_t* __restrict__ nbr,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ pin,
                                                  float* __restrict__ pout, int nact) {
    const int t = 2 * gf::xcd_remap(blockIdx.x, gridDim.x) + (threadIdx.x >> 7);
    if (t >= nact) return;
    const Quad Q = quad_of();
    const size_t base = (size_t)t * GF_TILE_CELLS;
    const size_t idx = base + Q.l0;
    const uint32_t fw = ld_flags(flags, idx);
    float c[4], bv[4], out[4];
    ld4(pin + idx, c);
    ld4(b + idx, bv);
    Nb4<float> o;
    load_nb4(pin, base, Q, nbr + (size_t)t * 6, o);
    for (int k = 0; k < 4; ++k) {
        const int d = fl_diag(fw, k);
        float v = 0.0f;
        if (fl_fluid(fw, k) && d > 0) v = (nb_sum4(o, c, k) - bv[k]) / (float)d;
        out[k] = v;
    }
    st4(pout + idx, out);
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
        hipLaunchKernelGGL(k_jacobi, dim3((n_active + 1) / 2), dim3(GF_WG), 0, st, s->flags,
                           s->nbr, s->b, in, out, n_active);
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
    if (check_every > 64) check_every = 64;
    if (iters) *iters = 0;
    if (rel_residual) *rel_residual = 0.0;
    if (replacements) *replacements = 0;
    if (kernel_ms) for (int i = 0; i < 4; ++i) kernel_ms[i] = 0.0;
    hipStream_t st = (hipStream_t)stream;
    if (n_active == 0) return 0;

    PcgArgs a;
    a.flags = s->flags; a.nbr = s->nbr; a.b = s->b;
    a.x32 = s->x32; a.r32 = s->r32; a.p32 = s->p32; a.ap32 = s->ap32;
    a.x64 = s->x64; a.r64 = s->r64; a.p64 = s->p64; a.ap64 = s->ap64;
    a.pc = s->pc; a.part = s->partials; a.log = (PcgLog*)s->scalars;
    a.nact = n_active;
    const int pairs = (n_active + 1) / 2;
    a.G = pairs < PCG_GMAX ? pairs : PCG_GMAX;
    a.rhs_sign = (float)rhs_sign;
    a.tol2 = tol * tol;
    a.delta2 = 0.01;
    a.max_iter = max_iter;
    const dim3 grid(a.G), block(GF_WG);
    const bool mixed = precision == GF_PCG_MIXED;

    if (mixed)
        hipLaunchKernelGGL((k_pcg_setup<float, true>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((k_pcg_setup<double, false>), grid, block, 0, st, a);
    int L = 1;
    hipLaunchKernelGGL(k_pcg_fin, dim3(1), block, 0, st, a, L);
    GF_LAUNCH_CHECK("k_pcg_setup/fin");

    // Optional per-kernel timing of the iterations that did work: events
    // around every launch of a chunk, read back at the chunk's host sync.
    const bool timed = kernel_ms != nullptr;
    hipEvent_t ev[4 * 64 + 2];
    int nev = 0;
    hipError_t e;
    if (timed) {
        nev = 4 * check_every + 2;
        for (int i = 0; i < nev; ++i) {
            e = hipEventCreate(&ev[i]);
            if (e != hipSuccess) {
                for (int j = 0; j < i; ++j) hipEventDestroy(ev[j]);
                return gf::set_error((int)e, "gf_scene_pcg_sync: event: %s", hipGetErrorString(e));
            }
        }
    }
    PcgLog hs;
    memset(&hs, 0, sizeof(hs));
    e = hipMemcpyAsync(&hs, &a.log[L & 1], sizeof(PcgLog), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = gf::set_error((int)e, "gf_scene_pcg_sync: readback: %s", hipGetErrorString(e));
    int iters_before = 0, nupd_before = 0, launched = 0;
    while (!rc && !hs.done) {
        const int chunk = check_every;
        for (int j = 0; j < chunk; ++j) {
            if (timed) hipEventRecord(ev[4 * j], st);
            ++L;
            if (mixed) hipLaunchKernelGGL(k_pcg_spmv<float>, grid, block, 0, st, a, L);
            else hipLaunchKernelGGL(k_pcg_spmv<double>, grid, block, 0, st, a, L);
            if (timed) hipEventRecord(ev[4 * j + 1], st);
            ++L;
            if (mixed) hipLaunchKernelGGL((k_pcg_update<float, true>), grid, block, 0, st, a, L);
            else hipLaunchKernelGGL((k_pcg_update<double, false>), grid, block, 0, st, a, L);
            if (timed) hipEventRecord(ev[4 * j + 2], st);
            ++L;
            if (mixed) hipLaunchKernelGGL((k_pcg_dir<float, true>), grid, block, 0, st, a, L);
            else hipLaunchKernelGGL((k_pcg_dir<double, false>), grid, block, 0, st, a, L);
            if (timed) hipEventRecord(ev[4 * j + 3], st);
        }
        launched += chunk;
        if (mixed) {
            if (timed) hipEventRecord(ev[4 * chunk], st);
            ++L;
            hipLaunchKernelGGL(k_pcg_replace, grid, block, 0, st, a, L);
            ++L;
            hipLaunchKernelGGL(k_pcg_fin, dim3(1), block, 0, st, a, L);
            if (timed) hipEventRecord(ev[4 * chunk + 1], st);
        }
        rc = gf::check_launch("pcg iteration");
        if (rc) break;
        e = hipMemcpyAsync(&hs, &a.log[L & 1], sizeof(PcgLog), hipMemcpyDeviceToHost, st);
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
