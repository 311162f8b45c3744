// Particle <-> grid transfers over 8^3 tiles of cell-sorted particles.
//
// P2G: one workgroup per tile owns the MAC faces of the tile's cells (plus the
// +1 plane needed for the tile's divergence), reads the particles of the tile
// and of its 1-cell halo (contiguous k-rows of the cell-sorted arrays), and
// accumulates the trilinear contributions in LDS.  No global float atomics:
// every face is finalised by exactly one workgroup, so normalize, gravity,
// the solid boundary condition and the divergence are fused into the same
// kernel.  Arithmetic follows solver3d.py:72-156 (FLIP), 1475-1588 (APIC),
// 758-799 (BC), 846-871 (divergence) and primitives/gridmath.py:71-95.
//
// G2P: one workgroup per tile stages the faces its particles can sample (and
// the cell pressures around them) in LDS, applies the pressure-gradient
// subtract and the BC there (solver3d.py:1364-1407, 758-799), writes the
// tile's own faces back, and gathers for its particles
// (solver3d.py:1414-1463 with primitives/gridmath.py:30-65, or 1591-1732).
#include "gf_common.h"

namespace {

struct P2GParams {
    const float* pos; const float* vel; const float* C;
    const int32_t* cell_start; const int32_t* marker; const int32_t* tile_map;
    float* u; float* v; float* w;
    float* us; float* vs; float* ws;
    float* uw; float* vw; float* ww;
    const float* su; const float* sv; const float* sw;
    float* b;
    const int32_t* tile_scene; const float* scene_par;  // packed scenes (NULL: one scene)
    const int32_t* tile_occ;  // staged P2G: bit 1 = the tile holds a particle, bit 2 = a solid
    const uint8_t* flags;     // staged P2G finish: compact flags (bit 3 = fluid)
    int nx, ny, nz, tnx, tny, tnz;
    float dx, g, dt, rho;
};

// Face region a tile finalises in P2G: component c spans 9 entries along its
// own axis (the +1 plane) and 8 along the others.
template <int COMP> struct P2GDims;
template <> struct P2GDims<0> { static constexpr int DI = 9, DJ = 8, DK = 8; };
template <> struct P2GDims<1> { static constexpr int DI = 8, DJ = 9, DK = 8; };
template <> struct P2GDims<2> { static constexpr int DI = 8, DJ = 8, DK = 9; };

constexpr int P2G_REGION = 576;

// Face region the tile's own particles reach (staged P2G): 9 entries along
// the component's own axis from the tile origin, 10 along the others from
// one below it.
template <int COMP> struct OwnDims;
template <> struct OwnDims<0> { static constexpr int DI = 9, DJ = 10, DK = 10; };
template <> struct OwnDims<1> { static constexpr int DI = 10, DJ = 9, DK = 10; };
template <> struct OwnDims<2> { static constexpr int DI = 10, DJ = 10, DK = 9; };
constexpr int OWN_REGION = 900;
constexpr int STAGE_PER_TILE = 3 * 2 * OWN_REGION;  // val and wt per component
#ifndef P2G_BATCH
#define P2G_BATCH 4
#endif

// Trilinear scatter of one velocity component into the tile's LDS region
// with LDS atomics (gridmath.py:71-95; out-of-grid corners dropped; corners
// outside the tile's region are finalised by another tile and skipped).
// Used only for particles whose corners fall outside their cell's stencil
// (particles outside the domain).
template <int COMP, bool APIC, class D>
__device__ __forceinline__ void scatter_comp(float* s_val, float* s_wt, float px, float py,
                                             float pz, float qx, float qy, float qz,
                                             float vp, const float* crow, int i0t,
                                             int j0t, int k0t, int nfi, int nfj,
                                             int nfk, float dx) {
    float cx = qx, cy = qy, cz = qz;
    if (COMP != 0) cx = qx - 0.5f;
    if (COMP != 1) cy = qy - 0.5f;
    if (COMP != 2) cz = qz - 0.5f;
    const int i0 = (int)floorf(cx), j0 = (int)floorf(cy), k0 = (int)floorf(cz);
    const float sx = cx - (float)i0, sy = cy - (float)j0, sz = cz - (float)k0;
#pragma unroll
    for (int di = 0; di < 2; ++di) {
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0 || ii >= nfi || jj < 0 || jj >= nfj || kk < 0 || kk >= nfk) continue;
                const int li = ii - i0t, lj = jj - j0t, lk = kk - k0t;
                if ((unsigned)li >= (unsigned)D::DI || (unsigned)lj >= (unsigned)D::DJ ||
                    (unsigned)lk >= (unsigned)D::DK)
                    continue;
                const float wxi = di ? sx : 1.0f - sx;
                const float wyi = dj ? sy : 1.0f - sy;
                const float wzi = dk ? sz : 1.0f - sz;
                const float wt = wxi * wyi * wzi;
                float val = vp;
                if (APIC) {
                    const float ox = (COMP == 0 ? (float)ii : (float)ii + 0.5f) * dx - px;
                    const float oy = (COMP == 1 ? (float)jj : (float)jj + 0.5f) * dx - py;
                    const float oz = (COMP == 2 ? (float)kk : (float)kk + 0.5f) * dx - pz;
                    val = vp + crow[0] * ox + crow[1] * oy + crow[2] * oz;
                }
                const int idx = (li * D::DJ + lj) * D::DK + lk;
                atomicAdd(&s_val[idx], val * wt);
                atomicAdd(&s_wt[idx], wt);
            }
        }
    }
}

// Per-cell stencil accumulator of one velocity component: the faces that
// particles of cell (ci,cj,ck) can reach are [ci, ci+1] along the
// component's own axis and [c-1, c+1] along the other two.
template <int COMP>
struct CellAcc {
    static constexpr int NX = COMP == 0 ? 2 : 3, NY = COMP == 1 ? 2 : 3, NZ = COMP == 2 ? 2 : 3;
    float val[NX][NY][NZ];
    float wt[NX][NY][NZ];
};

// Corner weights of one axis expanded onto the cell stencil: entries the
// particle does not touch get weight 0, touched entries get exactly the
// reference factor (1-s or s).  Returns false when the particle's corner
// origin lies outside the stencil.
template <int N>
__device__ __forceinline__ bool axis_weights(float c, int cell, float (&w)[N], float& s, int& org) {
    const float f0 = floorf(c);
    const int i0 = (int)f0;
    s = c - f0;
    const int o = i0 - (N == 2 ? cell : cell - 1);
    org = N == 2 ? cell : cell - 1;
    if (N == 2) {
        w[0] = 1.0f - s;
        w[1] = s;
        return o == 0;
    }
    w[0] = o == 0 ? 1.0f - s : 0.0f;
    w[1] = o == 0 ? s : 1.0f - s;
    w[N - 1] = o == 1 ? s : 0.0f;
    return (unsigned)o <= 1u;
}

template <int COMP, bool APIC>
__device__ __forceinline__ bool accumulate(CellAcc<COMP>& A, int ci, int cj, int ck, float px,
                                           float py, float pz, float qx, float qy, float qz,
                                           float vp, const float* crow, float dx) {
    using C = CellAcc<COMP>;
    float wx[C::NX], wy[C::NY], wz[C::NZ], sx, sy, sz;
    int ox0, oy0, oz0;
    bool ok = axis_weights<C::NX>(COMP == 0 ? qx : qx - 0.5f, ci, wx, sx, ox0);
    ok &= axis_weights<C::NY>(COMP == 1 ? qy : qy - 0.5f, cj, wy, sy, oy0);
    ok &= axis_weights<C::NZ>(COMP == 2 ? qz : qz - 0.5f, ck, wz, sz, oz0);
    if (!ok) return false;
    float offx[C::NX], offy[C::NY], offz[C::NZ];
    if (APIC) {
#pragma unroll
        for (int a = 0; a < C::NX; ++a)
            offx[a] = (COMP == 0 ? (float)(ox0 + a) : (float)(ox0 + a) + 0.5f) * dx - px;
#pragma unroll
        for (int b = 0; b < C::NY; ++b)
            offy[b] = (COMP == 1 ? (float)(oy0 + b) : (float)(oy0 + b) + 0.5f) * dx - py;
#pragma unroll
        for (int c = 0; c < C::NZ; ++c)
            offz[c] = (COMP == 2 ? (float)(oz0 + c) : (float)(oz0 + c) + 0.5f) * dx - pz;
    }
#pragma unroll
    for (int a = 0; a < C::NX; ++a)
#pragma unroll
        for (int b = 0; b < C::NY; ++b)
#pragma unroll
            for (int c = 0; c < C::NZ; ++c) {
                const float w = wx[a] * wy[b] * wz[c];
                float val = vp;
                if (APIC) val = vp + crow[0] * offx[a] + crow[1] * offy[b] + crow[2] * offz[c];
                A.val[a][b][c] += val * w;
                A.wt[a][b][c] += w;
            }
    return true;
}

// Adds a cell's stencil into the tile's LDS face region (plain read-modify-
// write; the caller guarantees no concurrent writer of the same entries).
template <int COMP, class D>
__device__ __forceinline__ void flush(const CellAcc<COMP>& A, float* s_val, float* s_wt, int ci,
                                      int cj, int ck, int i0t, int j0t, int k0t, int nx, int ny,
                                      int nz) {
    using C = CellAcc<COMP>;
    const int nfi = nx + (COMP == 0), nfj = ny + (COMP == 1), nfk = nz + (COMP == 2);
    const int fi0 = COMP == 0 ? ci : ci - 1, fj0 = COMP == 1 ? cj : cj - 1,
              fk0 = COMP == 2 ? ck : ck - 1;
    const int li0 = fi0 - i0t, lj0 = fj0 - j0t, lk0 = fk0 - k0t;
    if (li0 >= 0 && lj0 >= 0 && lk0 >= 0 && li0 + C::NX <= D::DI && lj0 + C::NY <= D::DJ &&
        lk0 + C::NZ <= D::DK && fi0 >= 0 && fj0 >= 0 && fk0 >= 0 && fi0 + C::NX <= nfi &&
        fj0 + C::NY <= nfj && fk0 + C::NZ <= nfk) {
        // whole stencil inside the region and the grid
#pragma unroll
        for (int a = 0; a < C::NX; ++a)
#pragma unroll
            for (int b = 0; b < C::NY; ++b)
#pragma unroll
                for (int c = 0; c < C::NZ; ++c) {
                    const int idx = ((li0 + a) * D::DJ + lj0 + b) * D::DK + lk0 + c;
                    s_val[idx] += A.val[a][b][c];
                    s_wt[idx] += A.wt[a][b][c];
                }
        return;
    }
#pragma unroll
    for (int a = 0; a < C::NX; ++a)
#pragma unroll
        for (int b = 0; b < C::NY; ++b)
#pragma unroll
            for (int c = 0; c < C::NZ; ++c) {
                const int fi = fi0 + a, fj = fj0 + b, fk = fk0 + c;
                const int li = fi - i0t, lj = fj - j0t, lk = fk - k0t;
                if (fi < 0 || fi >= nfi || fj < 0 || fj >= nfj || fk < 0 || fk >= nfk) continue;
                if ((unsigned)li >= (unsigned)D::DI || (unsigned)lj >= (unsigned)D::DJ ||
                    (unsigned)lk >= (unsigned)D::DK)
                    continue;
                const int idx = (li * D::DJ + lj) * D::DK + lk;
                s_val[idx] += A.val[a][b][c];
                s_wt[idx] += A.wt[a][b][c];
            }
}

// Flattened iteration over k-row segments of the cell-sorted particle arrays.
struct RowSet {
    int nrows, total;
};

__device__ __forceinline__ int row_find(const int* s_pref, int nrows, int t) {
    int lo = 0, hi = nrows - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (s_pref[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Rows (ci, cj) for ci in [ilo,ihi), cj in [jlo,jhi); each row covers cells
// k in [klo,khi).  Thread r < nrows computes row r; returns the flattened count.
__device__ __forceinline__ RowSet build_rows(const int32_t* __restrict__ cell_start, int ny,
                                             int nz, int ilo, int ihi, int jlo, int jhi,
                                             int klo, int khi, int* s_start, int* s_pref,
                                             int* s_wave) {
    const int nrj = jhi - jlo;
    const int nrows = (ihi > ilo && jhi > jlo && khi > klo) ? (ihi - ilo) * nrj : 0;
    int len = 0;
    if ((int)threadIdx.x < nrows) {
        const int ci = ilo + threadIdx.x / nrj, cj = jlo + threadIdx.x % nrj;
        const int base = (ci * ny + cj) * nz;
        const int s = cell_start[base + klo];
        const int e = cell_start[base + khi];
        s_start[threadIdx.x] = s;
        len = e - s;
    }
    int total;
    int ex = gf::wg_exclusive_scan(len, s_wave, &total);
    if ((int)threadIdx.x < nrows) s_pref[threadIdx.x] = ex;
    __syncthreads();
    RowSet r;
    r.nrows = nrows;
    r.total = total;
    return r;
}

// One component of the lane-per-cell P2G: accumulate the cell's stencil
// over particles [p0, p1) in registers, then add it to the LDS region in the
// flush phases r_lo..r_hi (phase = cell class; cls < 0 for empty cells).
// Cells and the region origin come in packed-grid indices; the arithmetic
// runs in the tile's scene frame F (its dx, its own cell indices and size).
template <int COMP, bool APIC, class D = P2GDims<COMP>>
__device__ __forceinline__ void p2g_component(const P2GParams& a, const gf::Frame& F,
                                              float* s_val, float* s_wt, int ci, int cj, int ck,
                                              int p0, int p1, int cls, int r_lo, int r_hi,
                                              int i0, int j0, int k0) {
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    const float dx = F.dx;
    ci -= F.ox; cj -= F.oy; ck -= F.oz;
    i0 -= F.ox; j0 -= F.oy; k0 -= F.oz;
    CellAcc<COMP> A;
#pragma unroll
    for (int q = 0; q < 18; ++q) {
        (&A.val[0][0][0])[q] = 0.0f;
        (&A.wt[0][0][0])[q] = 0.0f;
    }
    // The cell's particles go through in batches of P2G_BATCH: a batch's
    // loads are all issued before its math, so several are in flight.
    for (int pb = p0; pb < p1; pb += P2G_BATCH) {
        float bx[P2G_BATCH], by[P2G_BATCH], bz[P2G_BATCH], bv[P2G_BATCH];
        float bc[APIC ? P2G_BATCH : 1][3];
#pragma unroll
        for (int b = 0; b < P2G_BATCH; ++b) {
            const int p = pb + b < p1 ? pb + b : p1 - 1;
            bx[b] = a.pos[3 * p + 0]; by[b] = a.pos[3 * p + 1]; bz[b] = a.pos[3 * p + 2];
            bv[b] = a.vel[3 * p + COMP];
            if (APIC) {
                bc[APIC ? b : 0][0] = a.C[9 * p + 3 * COMP + 0];
                bc[APIC ? b : 0][1] = a.C[9 * p + 3 * COMP + 1];
                bc[APIC ? b : 0][2] = a.C[9 * p + 3 * COMP + 2];
            }
        }
#pragma unroll
        for (int b = 0; b < P2G_BATCH; ++b) {
            if (pb + b >= p1) break;
            const float px = bx[b], py = by[b], pz = bz[b], vp = bv[b];
            const float* crow = bc[APIC ? b : 0];
            const float qx = px / dx, qy = py / dx, qz = pz / dx;
            // corners outside the cell stencil (particles outside the domain)
            // are scattered with LDS atomics
            if (!accumulate<COMP, APIC>(A, ci, cj, ck, px, py, pz, qx, qy, qz, vp, crow, dx))
                scatter_comp<COMP, APIC, D>(s_val, s_wt, px, py, pz, qx, qy, qz, vp, crow, i0,
                                            j0, k0, nx + (COMP == 0), ny + (COMP == 1),
                                            nz + (COMP == 2), dx);
        }
    }
    for (int ph = r_lo; ph <= r_hi; ++ph) {
        __syncthreads();
        if (cls == ph) flush<COMP, D>(A, s_val, s_wt, ci, cj, ck, i0, j0, k0, nx, ny, nz);
    }
    __syncthreads();
}

template <bool APIC, bool RAW>
__global__ __launch_bounds__(GF_WG) void k_p2g_tile(P2GParams a) {
    __shared__ float s_val[3][P2G_REGION];
    __shared__ float s_wt[3][P2G_REGION];
    __shared__ int8_t s_mk[10][10][10];
    __shared__ int s_cls[28];
    __shared__ int s_start[128], s_pref[128];
    __shared__ int s_wave[GF_WG / GF_WAVE + 1];

    const int ntiles = a.tnx * a.tny * a.tnz;
    const int t = gf::xcd_remap(blockIdx.x, ntiles);
    const int tk = t % a.tnz, tj = (t / a.tnz) % a.tny, ti = t / (a.tnz * a.tny);
    const int i0 = ti * 8, j0 = tj * 8, k0 = tk * 8;
    const int nx = a.nx, ny = a.ny, nz = a.nz;

    for (int i = threadIdx.x; i < 3 * P2G_REGION; i += GF_WG) {
        (&s_val[0][0])[i] = 0.0f;
        (&s_wt[0][0])[i] = 0.0f;
    }
    if (!RAW) {
        for (int i = threadIdx.x; i < 1000; i += GF_WG) {
            const int ca = i / 100, cb = (i / 10) % 10, cc = i % 10;
            const int ci = i0 - 1 + ca, cj = j0 - 1 + cb, ck = k0 - 1 + cc;
            int8_t m = 0;
            if (ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz)
                m = (int8_t)a.marker[(ci * ny + cj) * nz + ck];
            s_mk[ca][cb][cc] = m;
        }
    }
    __syncthreads();

    // Lane per cell of the 1-cell-haloed tile: the lane runs through its
    // cell's particles (contiguous in the sorted arrays) and accumulates the
    // cell's face stencil in registers; stencils are then added to the LDS
    // region in 27 phases by cell class (ci%3, cj%3, ck%3) -- cells of one
    // class have disjoint stencils, so no atomics and a fixed summation order.
    const int ilo = max(i0 - 1, 0), jlo = max(j0 - 1, 0), klo = max(k0 - 1, 0);
    const int ni = min(i0 + 9, nx) - ilo, nj = min(j0 + 9, ny) - jlo, nk = min(k0 + 9, nz) - klo;
    const int ncell = ni * nj * nk;
    const float dx = a.dx;
    // Cells are enumerated class-major (class = relative coordinates mod 3),
    // so one flush phase involves one or two waves.  s_cls[r] = first linear
    // index of class r.
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int r = 0; r < 27; ++r) {
            const int ri = r / 9, rj = (r / 3) % 3, rk = r % 3;
            s_cls[r] = acc;
            acc += max(0, (ni - ri + 2) / 3) * max(0, (nj - rj + 2) / 3) * max(0, (nk - rk + 2) / 3);
        }
        s_cls[27] = acc;
    }
    // A tile whose haloed region holds no particle has nothing to scatter.
    const RowSet rs = build_rows(a.cell_start, ny, nz, ilo, ilo + ni, jlo, jlo + nj, klo, klo + nk,
                                 s_start, s_pref, s_wave);
    const int nrounds = rs.total > 0 ? (ncell + GF_WG - 1) / GF_WG : 0;
    for (int round = 0; round < nrounds; ++round) {
        const int c = round * GF_WG + threadIdx.x;
        int ci = 0, cj = 0, ck = 0, cls = -1, p0 = 0, p1 = 0;
        if (c < ncell) {
            int r = 0;
            while (r < 26 && s_cls[r + 1] <= c) ++r;
            const int ri = r / 9, rj = (r / 3) % 3, rk = r % 3;
            const int mj = max(0, (nj - rj + 2) / 3), mk = max(0, (nk - rk + 2) / 3);
            const int loc = c - s_cls[r];
            ck = klo + rk + 3 * (loc % mk);
            cj = jlo + rj + 3 * ((loc / mk) % mj);
            ci = ilo + ri + 3 * (loc / (mk * mj));
            const int key = (ci * ny + cj) * nz + ck;
            p0 = a.cell_start[key];
            p1 = a.cell_start[key + 1];
            if (p0 < p1) cls = r;
        }
        const int last = min(ncell, (round + 1) * GF_WG) - 1;
        int r_lo = 0, r_hi = 0;
        while (r_lo < 26 && s_cls[r_lo + 1] <= round * GF_WG) ++r_lo;
        while (r_hi < 26 && s_cls[r_hi + 1] <= last) ++r_hi;
        const gf::Frame F = gf::tile_frame(nullptr, nullptr, t, a.dx, nx, ny, nz);
        p2g_component<0, APIC>(a, F, s_val[0], s_wt[0], ci, cj, ck, p0, p1, cls, r_lo, r_hi, i0, j0, k0);
        p2g_component<1, APIC>(a, F, s_val[1], s_wt[1], ci, cj, ck, p0, p1, cls, r_lo, r_hi, i0, j0, k0);
        p2g_component<2, APIC>(a, F, s_val[2], s_wt[2], ci, cj, ck, p0, p1, cls, r_lo, r_hi, i0, j0, k0);
    }

    // Finalise the region: normalize (121-142), gravity (149-156), BC (758-799).
    const float gdt = a.g * a.dt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int DI = c == 0 ? 9 : 8, DJ = c == 1 ? 9 : 8, DK = c == 2 ? 9 : 8;
        const int fni = nx + (c == 0), fnj = ny + (c == 1), fnk = nz + (c == 2);
        float* gf_ = c == 0 ? a.u : (c == 1 ? a.v : a.w);
        for (int idx = threadIdx.x; idx < P2G_REGION; idx += GF_WG) {
            const int lk = idx % DK, lj = (idx / DK) % DJ, li = idx / (DK * DJ);
            const int ii = i0 + li, jj = j0 + lj, kk = k0 + lk;
            if (ii >= fni || jj >= fnj || kk >= fnk) continue;
            const bool owned = (c == 0 ? (li < 8 || ii == nx) : true) &&
                               (c == 1 ? (lj < 8 || jj == ny) : true) &&
                               (c == 2 ? (lk < 8 || kk == nz) : true);
            const size_t f = ((size_t)ii * fnj + jj) * fnk + kk;
            if (RAW) {
                if (owned) {
                    float* gw = c == 0 ? a.uw : (c == 1 ? a.vw : a.ww);
                    gf_[f] += s_val[c][idx];
                    gw[f] += s_wt[c][idx];
                }
                continue;
            }
            const float wt = s_wt[c][idx];
            float val = wt > 1.0e-8f ? s_val[c][idx] / wt : 0.0f;
            const float saved = val;
            if (c == 1) val = val + gdt;
            // Solid BC: cells on both sides of the face (local cell coords +1).
            const int ca = li + 1, cb = lj + 1, cc = lk + 1;
            bool lo_side, hi_side;
            const float* solid;
            if (c == 0) {
                lo_side = ii == 0 || s_mk[ca - 1][cb][cc] == 2;
                hi_side = ii == nx || s_mk[ca][cb][cc] == 2;
                solid = a.su;
            } else if (c == 1) {
                lo_side = jj == 0 || s_mk[ca][cb - 1][cc] == 2;
                hi_side = jj == ny || s_mk[ca][cb][cc] == 2;
                solid = a.sv;
            } else {
                lo_side = kk == 0 || s_mk[ca][cb][cc - 1] == 2;
                hi_side = kk == nz || s_mk[ca][cb][cc] == 2;
                solid = a.sw;
            }
            if (lo_side || hi_side) val = solid ? solid[f] : 0.0f;
            s_val[c][idx] = val;
            if (owned) {
                float* gs = c == 0 ? a.us : (c == 1 ? a.vs : a.ws);
                gf_[f] = val;
                gs[f] = saved;
            }
        }
    }
    if (RAW) return;
    __syncthreads();

    // Divergence of the tile's cells (846-871) into the compact rhs.
    const int slot = a.tile_map[t];
    if (slot < 0) return;
    const float scale = a.rho * a.dx / a.dt;
    const int lj = (threadIdx.x >> 3) & 7, lk = threadIdx.x & 7;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int li = (threadIdx.x >> 6) + 4 * h;
        const int l = (li * 8 + lj) * 8 + lk;
        float dv = 0.0f;
        if (i0 + li < nx && j0 + lj < ny && k0 + lk < nz && s_mk[li + 1][lj + 1][lk + 1] == 1) {
            const float u0 = s_val[0][(li * 8 + lj) * 8 + lk];
            const float u1 = s_val[0][((li + 1) * 8 + lj) * 8 + lk];
            const float v0 = s_val[1][(li * 9 + lj) * 8 + lk];
            const float v1 = s_val[1][(li * 9 + lj + 1) * 8 + lk];
            const float w0 = s_val[2][(li * 8 + lj) * 9 + lk];
            const float w1 = s_val[2][(li * 8 + lj) * 9 + lk + 1];
            const float d = u1 - u0 + v1 - v0 + w1 - w0;
            dv = d * scale;
        }
        a.b[(size_t)slot * GF_TILE_CELLS + l] = dv;
    }
}

// ------------------------------------------------ staged P2G (own cells) ----
// Two launches.  k_p2g_cells: each tile that holds particles scatters only
// the particles of its own 512 cells (each particle is read once, for all
// three components) into the face region those particles reach, and stores
// that region in `stage`.  k_p2g_finish: each tile sums, for every face it
// finalises, the staged partials of the tiles whose regions hold the face
// (at most two per axis) in a fixed order, then normalises, adds gravity,
// applies the BC and forms the divergence as k_p2g_tile does.
//
// k_p2g_cells, one workgroup of P2GC_NW waves per tile: wave w takes the
// i-planes [w*PPW, (w+1)*PPW) of the tile; in a plane lane (lj, lk) =
// (lane/8, lane%8) owns one cell and runs through the cell's particles (a
// contiguous run of the sorted arrays), accumulating every component's face
// stencil in registers.  A particle's position relative to its cell,
// t = p/dx - cell, is in [0,1] on each axis, so its stencil slots are fixed:
// on a component's own axis the faces cell and cell+1 with weights (1-t, t)
// (gridmath.py:71-95, fx = p/dx), on the other axes the faces cell-1, cell,
// cell+1 with weights (max(0.5-t,0), 1-|t-0.5|, max(t-0.5,0)), which is the
// reference's floor(f-0.5) corner pair with the unused slot at weight 0.
// The plane's stencils then go into the wave's own LDS face region: summed
// along k across lanes by DPP shifts (a cell's k-1/k+1 slots meet its
// neighbours' overlapping faces), then added along j in lj%3 phases (lj%2 on
// v's own axis), so no two lanes touch one entry in a phase and every
// entry's summation order is fixed.  The adds are plain LDS read-modify-
// writes: a wave's LDS accesses complete in program order, and LDS float
// atomics cost the kernel twice its time.  Waves then merge their regions in
// wave order.
// Particles outside the domain (t off [0,1] by more than P2GC_TOL, or NaN)
// skip the stencil and are scattered afterwards, one lane at a time, with
// the reference's own arithmetic (scatter of corners inside the grid).
// PACKED: tiles read their scene frame from the packed-scene table (dx,
// cell offset, size); cells are shifted to the scene's own indices.
constexpr int P2GC_NW = 2;                  // waves per workgroup
constexpr int P2GC_PPW = 8 / P2GC_NW;       // i-planes per wave
// One wave's face region: u i in [p0, p0+PPW], j,k in [-1, 8];
// v i in [p0-1, p0+PPW], j in [0, 8], k in [-1, 8];
// w i in [p0-1, p0+PPW], j in [-1, 8], k in [0, 8] (tile-relative faces).
constexpr int P2GC_NU = (P2GC_PPW + 1) * 100;
constexpr int P2GC_NV = (P2GC_PPW + 2) * 90;
constexpr int P2GC_NREG = P2GC_NU + 2 * P2GC_NV;
constexpr float P2GC_TOL = 1.0e-6f;


// Slots of component C's cell stencil: NA x NB x NC along x, y, z.
template <int C> struct Sten {
    static constexpr int NA = C == 0 ? 2 : 3, NB = C == 1 ? 2 : 3, NC = C == 2 ? 2 : 3;
};

template <int C>
__device__ __forceinline__ int p2gc_reg(int ri, int rj, int rk) {
    if (C == 0) return (ri * 10 + rj) * 10 + rk;
    if (C == 1) return P2GC_NU + (ri * 9 + rj) * 10 + rk;
    return P2GC_NU + P2GC_NV + (ri * 10 + rj) * 9 + rk;
}

// Lanes of one wave pass partial sums to each other through LDS by plain
// read-modify-writes in phases.  The hardware performs a wave's LDS accesses
// in program order; this fence (no instruction at wavefront scope) stops the
// compiler from moving or merging accesses of different phases, which from
// one lane's point of view are mutually exclusive branches.
__device__ __forceinline__ void p2gc_phase_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

template <int CTRL>
__device__ __forceinline__ float p2gc_dpp(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
enum { P2GC_ROW_SHL1 = 0x101, P2GC_ROW_SHR1 = 0x111 };

// Per-particle axis data: position relative to the cell, clamped to [0,1].
struct AxisT {
    float t[3];
    bool out;  // outside the domain: scattered by the exact path instead
};

__device__ __forceinline__ AxisT p2gc_axes(float x, float y, float z, float ihi, float ilo,
                                           float fci, float fcj, float fck) {
    AxisT r;
    // p/dx - cell with the quotient carried in two parts (1/dx = ihi + ilo),
    // one rounding at the size of t
    r.t[0] = fmaf(x, ilo, fmaf(x, ihi, -fci));
    r.t[1] = fmaf(y, ilo, fmaf(y, ihi, -fcj));
    r.t[2] = fmaf(z, ilo, fmaf(z, ihi, -fck));
    r.out = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        r.out |= !(fabsf(r.t[d] - 0.5f) <= 0.5f + P2GC_TOL);
        r.t[d] = fminf(fmaxf(r.t[d], 0.0f), 1.0f);
    }
    return r;
}

// Adds one particle to component C's stencil.  o: own-axis weights (already
// zero for a particle that must not count), g: staggered-axis weights, val:
// the value the weights multiply, per slot (FLIP: the velocity component).
template <int C>
__device__ __forceinline__ void p2gc_add(float (&S)[2][18], const float (&o)[3][2],
                                         const float (&g)[3][3], float v) {
    constexpr int A1 = C == 0 ? 1 : 0, A2 = C == 2 ? 1 : 2;  // the staggered axes
    float P[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n < 3; ++n) P[m][n] = g[A1][m] * g[A2][n];
    const float vo[2] = {v * o[C][0], v * o[C][1]};
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const int slot = C == 0 ? (s * 3 + m) * 3 + n
                                        : (C == 1 ? (m * 2 + s) * 3 + n : (m * 3 + n) * 2 + s);
                S[0][slot] = fmaf(vo[s], P[m][n], S[0][slot]);
                S[1][slot] = fmaf(o[C][s], P[m][n], S[1][slot]);
            }
}

// APIC: the slot value is v + c . (face position - particle position)
// (solver3d.py:1475-1588); fp: the stencil faces' world positions per axis.
template <int C>
__device__ __forceinline__ void p2gc_add_apic(float (&S)[2][18], const float (&o)[3][2],
                                              const float (&g)[3][3], float v, const float* c,
                                              float px, float py, float pz,
                                              const float (&fp)[3][3]) {
    using T = Sten<C>;
    float ox[3], oy[3], oz[3];
#pragma unroll
    for (int a = 0; a < T::NA; ++a) ox[a] = fp[0][a] - px;
#pragma unroll
    for (int b = 0; b < T::NB; ++b) oy[b] = fp[1][b] - py;
#pragma unroll
    for (int e = 0; e < T::NC; ++e) oz[e] = fp[2][e] - pz;
    const float* wa = C == 0 ? o[0] : g[0];
    const float* wb = C == 1 ? o[1] : g[1];
    const float* wc = C == 2 ? o[2] : g[2];
#pragma unroll
    for (int a = 0; a < T::NA; ++a) {
        const float va = fmaf(c[0], ox[a], v);
#pragma unroll
        for (int b = 0; b < T::NB; ++b) {
            const float wab = wa[a] * wb[b];
            const float vab = fmaf(c[1], oy[b], va);
#pragma unroll
            for (int e = 0; e < T::NC; ++e) {
                const int slot = (a * T::NB + b) * T::NC + e;
                const float w = wab * wc[e];
                S[0][slot] = fmaf(fmaf(c[2], oz[e], vab), w, S[0][slot]);
                S[1][slot] += w;
            }
        }
    }
}

__device__ __forceinline__ void p2gc_weights(const AxisT& r, bool keep, float (&o)[3][2],
                                             float (&g)[3][3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float t = r.t[d];
        o[d][0] = keep ? 1.0f - t : 0.0f;
        o[d][1] = keep ? t : 0.0f;
        const float h = t - 0.5f;
        g[d][0] = fmaxf(-h, 0.0f);
        g[d][1] = 1.0f - fabsf(h);
        g[d][2] = fmaxf(h, 0.0f);
    }
}

// Writes the plane's component-C stencils into the wave's region (see the
// section comment).  cl: the lane's cell in scene indices, sn: scene size.
template <int C>
__device__ __forceinline__ void p2gc_flush(float (&S)[2][18], float* reg, int lp, int lj,
                                           int lk, const int (&cl)[3], const int (&sn)[3]) {
    using T = Sten<C>;
    // entries of faces outside the scene's grid are dropped (staggered axes
    // at the first and last cell; own-axis faces always lie in the grid)
#pragma unroll
    for (int a = 0; a < T::NA; ++a)
#pragma unroll
        for (int b = 0; b < T::NB; ++b)
#pragma unroll
            for (int e = 0; e < T::NC; ++e) {
                bool drop = false;
                if (C != 0) drop |= (a == 0 && cl[0] == 0) || (a == 2 && cl[0] == sn[0] - 1);
                if (C != 1) drop |= (b == 0 && cl[1] == 0) || (b == 2 && cl[1] == sn[1] - 1);
                if (C != 2) drop |= (e == 0 && cl[2] == 0) || (e == 2 && cl[2] == sn[2] - 1);
                const int slot = (a * T::NB + b) * T::NC + e;
                if (drop) {
                    S[0][slot] = 0.0f;
                    S[1][slot] = 0.0f;
                }
            }
    // along k: face k (this lane's cell) takes this cell's middle slot, then
    // the k-1 cell's upper slot, then the k+1 cell's lower slot; the tile's
    // edge lanes keep the slots of the faces outside the tile's k range
    constexpr int NAB = T::NA * T::NB;
    float R[2][NAB], E[2][NAB];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int ab = 0; ab < NAB; ++ab) {
            if (T::NC == 3) {
                float lo = p2gc_dpp<P2GC_ROW_SHR1>(S[q][ab * 3 + 2]);
                float hi = p2gc_dpp<P2GC_ROW_SHL1>(S[q][ab * 3 + 0]);
                lo = lk == 0 ? 0.0f : lo;
                hi = lk == 7 ? 0.0f : hi;
                R[q][ab] = (S[q][ab * 3 + 1] + lo) + hi;
                E[q][ab] = lk == 0 ? S[q][ab * 3 + 0] : S[q][ab * 3 + 2];
            } else {
                float lo = p2gc_dpp<P2GC_ROW_SHR1>(S[q][ab * 2 + 1]);
                lo = lk == 0 ? 0.0f : lo;
                R[q][ab] = S[q][ab * 2 + 0] + lo;
                E[q][ab] = S[q][ab * 2 + 1];
            }
        }
    const int rk = T::NC == 3 ? lk + 1 : lk;
    const bool edge = T::NC == 3 ? (lk == 0 || lk == 7) : lk == 7;
    const int rke = T::NC == 3 ? (lk == 0 ? 0 : 9) : 8;
    constexpr int NPH = T::NB == 3 ? 3 : 2;
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph) {
        p2gc_phase_fence();
        if (lj % NPH != ph) continue;
#pragma unroll
        for (int a = 0; a < T::NA; ++a)
#pragma unroll
            for (int b = 0; b < T::NB; ++b) {
                const int idx = p2gc_reg<C>(lp + a, lj + b, rk);
                reg[idx] += R[0][a * T::NB + b];
                reg[P2GC_NREG + idx] += R[1][a * T::NB + b];
            }
        if (edge) {
#pragma unroll
            for (int a = 0; a < T::NA; ++a)
#pragma unroll
                for (int b = 0; b < T::NB; ++b) {
                    const int idx = p2gc_reg<C>(lp + a, lj + b, rke);
                    reg[idx] += E[0][a * T::NB + b];
                    reg[P2GC_NREG + idx] += E[1][a * T::NB + b];
                }
        }
    }
    p2gc_phase_fence();
}

// The reference's scatter (gridmath.py:71-95) of one particle's component C
// into the wave's region: f = p/dx (IEEE), corners floor(f - 0.5 off-axis),
// corners outside the scene's grid dropped.  o: the region's origin in scene
// face indices.
template <int C, bool APIC>
__device__ __forceinline__ void p2gc_exact(float* reg, float px, float py, float pz, float vp,
                                           const float* crow, float dx, const int (&sn)[3],
                                           const int (&o)[3]) {
    float cx = px / dx, cy = py / dx, cz = pz / dx;
    if (C != 0) cx -= 0.5f;
    if (C != 1) cy -= 0.5f;
    if (C != 2) cz -= 0.5f;
    const int i0 = (int)floorf(cx), j0 = (int)floorf(cy), k0 = (int)floorf(cz);
    const float sx = cx - (float)i0, sy = cy - (float)j0, sz = cz - (float)k0;
    const int nfi = sn[0] + (C == 0), nfj = sn[1] + (C == 1), nfk = sn[2] + (C == 2);
    const int RI = C == 0 ? P2GC_PPW + 1 : P2GC_PPW + 2, RJ = C == 1 ? 9 : 10,
              RK = C == 2 ? 9 : 10;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0 || ii >= nfi || jj < 0 || jj >= nfj || kk < 0 || kk >= nfk) continue;
                const int ri = ii - o[0], rj = jj - o[1], rk = kk - o[2];
                if ((unsigned)ri >= (unsigned)RI || (unsigned)rj >= (unsigned)RJ ||
                    (unsigned)rk >= (unsigned)RK)
                    continue;
                const float wt = (di ? sx : 1.0f - sx) * (dj ? sy : 1.0f - sy) * (dk ? sz : 1.0f - sz);
                float val = vp;
                if (APIC) {
                    const float ox = (C == 0 ? (float)ii : (float)ii + 0.5f) * dx - px;
                    const float oy = (C == 1 ? (float)jj : (float)jj + 0.5f) * dx - py;
                    const float oz = (C == 2 ? (float)kk : (float)kk + 0.5f) * dx - pz;
                    val = vp + crow[0] * ox + crow[1] * oy + crow[2] * oz;
                }
                const int idx = p2gc_reg<C>(ri, rj, rk);
                reg[idx] += val * wt;
                reg[P2GC_NREG + idx] += wt;
            }
}

// One particle's data for the components [C0, C1).
struct P2GCPart {
    float x, y, z, v[3], c[3][3];
};

template <bool APIC, int C0, int C1>
__device__ __forceinline__ void p2gc_load(const P2GParams& a, int q, P2GCPart& d) {
    d.x = a.pos[3 * q];
    d.y = a.pos[3 * q + 1];
    d.z = a.pos[3 * q + 2];
#pragma unroll
    for (int c = C0; c < C1; ++c) {
        d.v[c] = a.vel[3 * q + c];
        if (APIC) {
#pragma unroll
            for (int e = 0; e < 3; ++e) d.c[c][e] = a.C[9 * q + 3 * c + e];
        }
    }
}

// Accumulates the lane's cell particles [p0, p1) for the components
// [C0, C1) into S; cur holds particle p0 already.  The next particle is
// loaded before the current one is added, so its load overlaps the math.
template <bool APIC, int C0, int C1>
__device__ __forceinline__ void p2gc_accumulate(const P2GParams& a, float (&S)[3][2][18],
                                                P2GCPart cur, int p0, int p1, float ihi,
                                                float ilo, const int (&cl)[3], float dx,
                                                bool& anyout) {
    const float fci = (float)cl[0], fcj = (float)cl[1], fck = (float)cl[2];
    // APIC: world positions of the stencil faces per axis (own axis of
    // component C: cell, cell+1; others: cell-1, cell, cell+1, at +0.5)
    float fp[3][3][3];
    if (APIC) {
#pragma unroll
        for (int c = C0; c < C1; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    fp[c][d][m] = d == c ? (float)(cl[d] + m) * dx
                                         : ((float)(cl[d] - 1 + m) + 0.5f) * dx;
    }
    for (int p = p0; p < p1; ++p) {
        P2GCPart nxt;
        p2gc_load<APIC, C0, C1>(a, p + 1 < p1 ? p + 1 : p, nxt);
        const AxisT r = p2gc_axes(cur.x, cur.y, cur.z, ihi, ilo, fci, fcj, fck);
        anyout |= r.out;
        float o[3][2], g[3][3];
        p2gc_weights(r, !r.out, o, g);
#pragma unroll
        for (int c = C0; c < C1; ++c) {
            if (APIC) {
                if (c == 0) p2gc_add_apic<0>(S[0], o, g, cur.v[0], cur.c[0], cur.x, cur.y, cur.z, fp[0]);
                if (c == 1) p2gc_add_apic<1>(S[1], o, g, cur.v[1], cur.c[1], cur.x, cur.y, cur.z, fp[1]);
                if (c == 2) p2gc_add_apic<2>(S[2], o, g, cur.v[2], cur.c[2], cur.x, cur.y, cur.z, fp[2]);
            } else {
                if (c == 0) p2gc_add<0>(S[0], o, g, cur.v[0]);
                if (c == 1) p2gc_add<1>(S[1], o, g, cur.v[1]);
                if (c == 2) p2gc_add<2>(S[2], o, g, cur.v[2]);
            }
        }
        cur = nxt;
    }
}

// Particles outside the domain among the lane's [p0, p1), one lane at a time
// in lane order, with the reference's arithmetic.
template <bool APIC, int C0, int C1>
__device__ __forceinline__ void p2gc_outside(const P2GParams& a, float* reg, bool anyout, int p0,
                                             int p1, float ihi, float ilo, const int (&cl)[3],
                                             const int (&sn)[3], float dx,
                                             const int (&org)[3]) {
    const float fci = (float)cl[0], fcj = (float)cl[1], fck = (float)cl[2];
    unsigned long long m = __ballot(anyout);
    while (m) {
        const int L = __ffsll((long long)m) - 1;
        m &= m - 1;
        p2gc_phase_fence();
        if ((int)(threadIdx.x & 63) != L) continue;
        for (int p = p0; p < p1; ++p) {
            const float px = a.pos[3 * p], py = a.pos[3 * p + 1], pz = a.pos[3 * p + 2];
            if (!p2gc_axes(px, py, pz, ihi, ilo, fci, fcj, fck).out) continue;
#pragma unroll
            for (int c = C0; c < C1; ++c) {
                float crow[3] = {0.0f, 0.0f, 0.0f};
                if (APIC)
                    for (int e = 0; e < 3; ++e) crow[e] = a.C[9 * p + 3 * c + e];
                const float vp = a.vel[3 * p + c];
                // region origin of component c in scene face indices
                const int oc[3] = {org[0] - (c != 0), org[1] - (c != 1), org[2] - (c != 2)};
                if (c == 0) p2gc_exact<0, APIC>(reg, px, py, pz, vp, crow, dx, sn, oc);
                if (c == 1) p2gc_exact<1, APIC>(reg, px, py, pz, vp, crow, dx, sn, oc);
                if (c == 2) p2gc_exact<2, APIC>(reg, px, py, pz, vp, crow, dx, sn, oc);
            }
        }
    }
}

template <bool APIC, bool PACKED>
__device__ __forceinline__ void p2g_cells_body(const P2GParams& a, float* __restrict__ stage,
                                               int32_t* __restrict__ has) {
    __shared__ float s_reg[P2GC_NW][2 * P2GC_NREG];
    const int ntiles = a.tnx * a.tny * a.tnz;
    const int t = gf::xcd_remap(blockIdx.x, ntiles);
    const int occ = a.tile_occ[t];
    if (threadIdx.x == 0) has[t] = (occ >> 1) & 1;
    if (!(occ & 2)) return;
    const int tk = t % a.tnz, tj = (t / a.tnz) % a.tny, ti = t / (a.tnz * a.tny);
    const int i0 = ti * 8, j0 = tj * 8, k0 = tk * 8;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const gf::Frame F =
        gf::tile_frame(PACKED ? a.tile_scene : nullptr, a.scene_par, t, a.dx, nx, ny, nz);
    const float dx = F.dx;
    const float ihi = 1.0f / dx;
    const float ilo = fmaf(-ihi, dx, 1.0f) * ihi;
    const int sn[3] = {F.nx, F.ny, F.nz};
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lj = lane >> 3, lk = lane & 7;
    float* reg = s_reg[wv];
    for (int e = lane; e < 2 * P2GC_NREG; e += 64) reg[e] = 0.0f;
    // region origin (u faces; v/w one lower on their staggered axes) in
    // scene face indices
    const int org[3] = {i0 - F.ox + wv * P2GC_PPW, j0 - F.oy, k0 - F.oz};
    const int cj = j0 + lj, ck = k0 + lk;
    auto range = [&](int lp, int& q0, int& q1) {
        const int ci = i0 + wv * P2GC_PPW + lp;
        q0 = q1 = 0;
        if (ci < nx && cj < ny && ck < nz) {
            const int key = (ci * ny + cj) * nz + ck;
            q0 = a.cell_start[key];
            q1 = a.cell_start[key + 1];
        }
    };
    int p0, p1;
    range(0, p0, p1);
    P2GCPart cur;
    if (!APIC && p0 < p1) p2gc_load<false, 0, 3>(a, p0, cur);
    for (int lp = 0; lp < P2GC_PPW; ++lp) {
        int n0 = 0, n1 = 0;
        if (lp + 1 < P2GC_PPW) range(lp + 1, n0, n1);
        const int cl[3] = {i0 + wv * P2GC_PPW + lp - F.ox, cj - F.oy, ck - F.oz};
        float S[3][2][18];
        bool anyout = false;
        if (APIC) {
            // one component per pass keeps the affine rows in registers
#pragma unroll
            for (int c = 0; c < 3; ++c) {
#pragma unroll
                for (int q = 0; q < 18; ++q) S[c][0][q] = S[c][1][q] = 0.0f;
                P2GCPart first;
                bool out_c = false;
                if (c == 0) {
                    if (p0 < p1) p2gc_load<true, 0, 1>(a, p0, first);
                    p2gc_accumulate<true, 0, 1>(a, S, first, p0, p1, ihi, ilo, cl, dx, out_c);
                    p2gc_flush<0>(S[0], reg, lp, lj, lk, cl, sn);
                    p2gc_outside<true, 0, 1>(a, reg, out_c, p0, p1, ihi, ilo, cl, sn, dx, org);
                } else if (c == 1) {
                    if (p0 < p1) p2gc_load<true, 1, 2>(a, p0, first);
                    p2gc_accumulate<true, 1, 2>(a, S, first, p0, p1, ihi, ilo, cl, dx, out_c);
                    p2gc_flush<1>(S[1], reg, lp, lj, lk, cl, sn);
                    p2gc_outside<true, 1, 2>(a, reg, out_c, p0, p1, ihi, ilo, cl, sn, dx, org);
                } else {
                    if (p0 < p1) p2gc_load<true, 2, 3>(a, p0, first);
                    p2gc_accumulate<true, 2, 3>(a, S, first, p0, p1, ihi, ilo, cl, dx, out_c);
                    p2gc_flush<2>(S[2], reg, lp, lj, lk, cl, sn);
                    p2gc_outside<true, 2, 3>(a, reg, out_c, p0, p1, ihi, ilo, cl, sn, dx, org);
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int q = 0; q < 18; ++q) S[c][0][q] = S[c][1][q] = 0.0f;
            p2gc_accumulate<false, 0, 3>(a, S, cur, p0, p1, ihi, ilo, cl, dx, anyout);
            // the next plane's first particle is loaded while this plane's
            // stencils are written out
            if (lp + 1 < P2GC_PPW && n0 < n1) p2gc_load<false, 0, 3>(a, n0, cur);
            p2gc_flush<0>(S[0], reg, lp, lj, lk, cl, sn);
            p2gc_flush<1>(S[1], reg, lp, lj, lk, cl, sn);
            p2gc_flush<2>(S[2], reg, lp, lj, lk, cl, sn);
            p2gc_outside<false, 0, 3>(a, reg, anyout, p0, p1, ihi, ilo, cl, sn, dx, org);
        }
        p0 = n0;
        p1 = n1;
    }
    __syncthreads();
    // merge the waves' regions (in wave order) into the tile's staged region
    float* st = stage + (size_t)t * STAGE_PER_TILE;
    for (int e = threadIdx.x; e < 3 * OWN_REGION; e += 64 * P2GC_NW) {
        const int c = e / OWN_REGION, r = e - c * OWN_REGION;
        const int plane = c == 0 ? 100 : 90;
        const int si = r / plane, rest = r - si * plane;
        const int base = c == 0 ? 0 : (c == 1 ? P2GC_NU : P2GC_NU + P2GC_NV);
        const int rmax = c == 0 ? P2GC_PPW : P2GC_PPW + 1;
        float sv = 0.0f, sw = 0.0f;
#pragma unroll
        for (int w = 0; w < P2GC_NW; ++w) {
            const int ri = si - w * P2GC_PPW;
            if (ri >= 0 && ri <= rmax) {
                const int idx = base + ri * plane + rest;
                sv += s_reg[w][idx];
                sw += s_reg[w][P2GC_NREG + idx];
            }
        }
        st[c * 2 * OWN_REGION + r] = sv;
        st[c * 2 * OWN_REGION + OWN_REGION + r] = sw;
    }
}

template <bool PACKED>
__global__ __launch_bounds__(64 * P2GC_NW) void k_p2g_cells_flip(P2GParams a,
                                                                  float* __restrict__ stage,
                                                                  int32_t* __restrict__ has) {
    p2g_cells_body<false, PACKED>(a, stage, has);
}

template <bool PACKED>
__global__ __launch_bounds__(64 * P2GC_NW) void k_p2g_cells_apic(P2GParams a,
                                                                  float* __restrict__ stage,
                                                                  int32_t* __restrict__ has) {
    p2g_cells_body<true, PACKED>(a, stage, has);
}

// The marker box is loaded only when a neighbouring tile holds a solid cell
// (tile_occ bit 2); otherwise the BC applies at the grid boundary alone.  The
// divergence takes the fluid test from the compact flags (bit 3 = marker 1).
__global__ __launch_bounds__(GF_WG) void k_p2g_finish(P2GParams a,
                                                      const float* __restrict__ stage,
                                                      const int32_t* __restrict__ has) {
    __shared__ float s_val[3][P2G_REGION];
    __shared__ int8_t s_mk[10][10][10];
    __shared__ int s_has[27];

    const int ntiles = a.tnx * a.tny * a.tnz;
    const int t = gf::xcd_remap(blockIdx.x, ntiles);
    const int tk = t % a.tnz, tj = (t / a.tnz) % a.tny, ti = t / (a.tnz * a.tny);
    const int i0 = ti * 8, j0 = tj * 8, k0 = tk * 8;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    int solid_nb = 0;
    if (threadIdx.x < 27) {
        const int di = (int)threadIdx.x / 9 - 1, dj = ((int)threadIdx.x / 3) % 3 - 1,
                  dk = (int)threadIdx.x % 3 - 1;
        const int a_ = ti + di, b_ = tj + dj, c_ = tk + dk;
        int h = 0;
        if (a_ >= 0 && a_ < a.tnx && b_ >= 0 && b_ < a.tny && c_ >= 0 && c_ < a.tnz) {
            const int q = (a_ * a.tny + b_) * a.tnz + c_;
            h = has[q];
            solid_nb = (a.tile_occ[q] >> 2) & 1;
        }
        s_has[threadIdx.x] = h;
    }
    const bool any_solid = __syncthreads_or(solid_nb);
    if (any_solid) {
        for (int i = threadIdx.x; i < 1000; i += GF_WG) {
            const int ca = i / 100, cb = (i / 10) % 10, cc = i % 10;
            const int ci = i0 - 1 + ca, cj = j0 - 1 + cb, ck = k0 - 1 + cc;
            int8_t m = 0;
            if (ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz)
                m = (int8_t)a.marker[(ci * ny + cj) * nz + ck];
            s_mk[ca][cb][cc] = m;
        }
        __syncthreads();
    }
    bool any_has = false;
    for (int q = 0; q < 27; ++q) any_has |= s_has[q] != 0;

    const float gdt =
        a.tile_scene ? a.scene_par[gf::SCENE_PAR * max(a.tile_scene[t], 0)] : a.g * a.dt;
    const int slot = a.tile_map[t];
    // A tile away from the grid boundary with no particle and no solid cell in
    // its 27-tile neighbourhood (so no fluid cell either) gets what the face
    // loop below computes for it: no weight, so u = w = 0, v = g dt (gravity
    // added to 0), and the saved pre-gravity faces 0; no BC applies.
    if (!any_has && !any_solid && slot < 0 && ti > 0 && ti < a.tnx - 1 && tj > 0 &&
        tj < a.tny - 1 && tk > 0 && tk < a.tnz - 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int fnj = ny + (c == 1), fnk = nz + (c == 2);
            float* gf_ = c == 0 ? a.u : (c == 1 ? a.v : a.w);
            float* gs = c == 0 ? a.us : (c == 1 ? a.vs : a.ws);
            const float val = c == 1 ? gdt : 0.0f;
            for (int e = threadIdx.x; e < GF_TILE_CELLS; e += GF_WG) {
                const int li = e >> 6, lj = (e >> 3) & 7, lk = e & 7;
                const size_t f = ((size_t)(i0 + li) * fnj + (j0 + lj)) * fnk + (k0 + lk);
                gf_[f] = val;
                gs[f] = 0.0f;
            }
        }
        return;
    }
    const gf::Frame F = gf::tile_frame(a.tile_scene, a.scene_par, t, a.dx, nx, ny, nz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int DJ = c == 1 ? 9 : 8, DK = c == 2 ? 9 : 8;
        const int RJ = c == 1 ? 9 : 10, RK = c == 2 ? 9 : 10;  // staged region dims
        const int fni = nx + (c == 0), fnj = ny + (c == 1), fnk = nz + (c == 2);
        float* gf_ = c == 0 ? a.u : (c == 1 ? a.v : a.w);
        // staged-region strides: one tile along x, y, z (in floats)
        const ptrdiff_t sk = STAGE_PER_TILE, sj = sk * a.tnz, si = sj * a.tny;
        const float* stage_c = stage + (size_t)t * STAGE_PER_TILE + (size_t)c * 2 * OWN_REGION;
        // a thread's (at most three) faces of the component's region, unrolled
#pragma unroll
        for (int r = 0; r < (P2G_REGION + GF_WG - 1) / GF_WG; ++r) {
            const int idx = (int)threadIdx.x + r * GF_WG;
            if (idx >= P2G_REGION) continue;
            const int lk = idx % DK, lj = (idx / DK) % DJ, li = idx / (DK * DJ);
            const int ii = i0 + li, jj = j0 + lj, kk = k0 + lk;
            if (ii >= fni || jj >= fnj || kk >= fnk) continue;
            const bool owned = (c == 0 ? (li < 8 || ii == nx) : true) &&
                               (c == 1 ? (lj < 8 || jj == ny) : true) &&
                               (c == 2 ? (lk < 8 || kk == nz) : true);
            const size_t f = ((size_t)ii * fnj + jj) * fnk + kk;
            // partial sums of the tiles whose staged regions hold this face:
            // offsets d in [lo, hi] per axis, ascending (-1 only at local
            // coordinate 0, +1 only at the last one)
            float sv = 0.0f, sw = 0.0f;
            const int ilo = li == 0 ? -1 : 0, ihi = li == (c == 0 ? 8 : 7) ? 1 : 0;
            const int jlo = lj == 0 ? -1 : 0, jhi = lj == (c == 1 ? 8 : 7) ? 1 : 0;
            const int klo = lk == 0 ? -1 : 0, khi = lk == (c == 2 ? 8 : 7) ? 1 : 0;
            // this tile's entry (region origin one below its corner on the
            // non-own axes); a neighbour (di,dj,dk) holds the face at the
            // same entry shifted by one tile and by -8 entries per axis
            const float* own = stage_c + ((li + (c != 0)) * RJ + lj + (c != 1)) * RK + lk +
                               (c != 2);
            if (any_has) {
                // the (at most 2 x 2 x 2) offsets in ascending order, per axis
                // d = lo + b for b < n (n = 2 at a region border).  Every
                // load is issued (an absent one reads this tile's own entry,
                // always in the buffer) and masked to +0 afterwards, so all
                // are in flight together; adding +0 leaves the partial sum
                // (never -0) unchanged, so the sum is the one over the
                // present offsets in the same order.
                const int ni = 1 + ((ilo | ihi) != 0), nj = 1 + ((jlo | jhi) != 0),
                          nk = 1 + ((klo | khi) != 0);
                float cv[8], cw[8];
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    const int bi = m >> 2, bj = (m >> 1) & 1, bk = m & 1;
                    const int di = ilo + bi, dj = jlo + bj, dk = klo + bk;
                    const bool ok = bi < ni && bj < nj && bk < nk &&
                                    s_has[(di + 1) * 9 + (dj + 1) * 3 + (dk + 1)];
                    const ptrdiff_t off = (ptrdiff_t)di * (si - 8 * RJ * RK) +
                                          (ptrdiff_t)dj * (sj - 8 * RK) + (ptrdiff_t)dk * (sk - 8);
                    const float* st = own + (ok ? off : 0);
                    const uint32_t keep = 0u - (uint32_t)ok;
                    cv[m] = __uint_as_float(__float_as_uint(st[0]) & keep);
                    cw[m] = __uint_as_float(__float_as_uint(st[OWN_REGION]) & keep);
                }
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    sv += cv[m];
                    sw += cw[m];
                }
            }
            float val = sw > 1.0e-8f ? sv / sw : 0.0f;
            const float saved = val;
            if (c == 1) val = val + gdt;
            const int ca = li + 1, cb = lj + 1, cc = lk + 1;
            bool lo_side, hi_side;
            const float* solid;
            if (c == 0) {
                lo_side = ii == 0 || (any_solid && s_mk[ca - 1][cb][cc] == 2);
                hi_side = ii == nx || (any_solid && s_mk[ca][cb][cc] == 2);
                solid = a.su;
            } else if (c == 1) {
                lo_side = jj == 0 || (any_solid && s_mk[ca][cb - 1][cc] == 2);
                hi_side = jj == ny || (any_solid && s_mk[ca][cb][cc] == 2);
                solid = a.sv;
            } else {
                lo_side = kk == 0 || (any_solid && s_mk[ca][cb][cc - 1] == 2);
                hi_side = kk == nz || (any_solid && s_mk[ca][cb][cc] == 2);
                solid = a.sw;
            }
            if (lo_side || hi_side) val = solid ? solid[f] : 0.0f;
            s_val[c][idx] = val;
            if (owned) {
                float* gs = c == 0 ? a.us : (c == 1 ? a.vs : a.ws);
                gf_[f] = val;
                gs[f] = saved;
            }
        }
    }
    if (slot < 0) return;
    __syncthreads();

    const float scale = a.rho * F.dx / a.dt;
    const int lj = (threadIdx.x >> 3) & 7, lk = threadIdx.x & 7;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int li = (threadIdx.x >> 6) + 4 * h;
        const int l = (li * 8 + lj) * 8 + lk;
        float dv = 0.0f;
        if (i0 + li < nx && j0 + lj < ny && k0 + lk < nz &&
            (a.flags[(size_t)slot * GF_TILE_CELLS + l] & 8)) {
            const float u0 = s_val[0][(li * 8 + lj) * 8 + lk];
            const float u1 = s_val[0][((li + 1) * 8 + lj) * 8 + lk];
            const float v0 = s_val[1][(li * 9 + lj) * 8 + lk];
            const float v1 = s_val[1][(li * 9 + lj + 1) * 8 + lk];
            const float w0 = s_val[2][(li * 8 + lj) * 9 + lk];
            const float w1 = s_val[2][(li * 8 + lj) * 9 + lk + 1];
            const float d = u1 - u0 + v1 - v0 + w1 - w0;
            dv = d * scale;
        }
        a.b[(size_t)slot * GF_TILE_CELLS + l] = dv;
    }
}

// ---------------------------------------------------------------- G2P ----

struct G2PParams {
    float* pos; float* vel; float* C;
    const int32_t* cell_start; const int32_t* marker; const int32_t* tile_map;
    float* u; float* v; float* w;
    const float* us; const float* vs; const float* ws;
    const float* su; const float* sv; const float* sw;
    const float* pc;
    // Pre-projection faces; NULL: rebuilt from us/vs/ws as P2G produced them.
    const float* pu; const float* pv; const float* pw;
    const int32_t* tile_scene; const float* scene_par;  // packed scenes (NULL: one scene)
    int nx, ny, nz, tnx, tny, tnz;
    float dx, dt, rho, flip_blend, gdt;
    float domx, domy, domz;
};

// Stands in for an absent value in loads that are issued unconditionally.
__device__ const float g_zero_f = 0.0f;

// Face region a tile samples: component c spans [o, o+9) on its own axis with
// origin at the tile corner and [o-1, o+9) (10 entries) on the other axes.
constexpr int G2P_REGION = 900;

template <int COMP>
__device__ __forceinline__ int g2p_idx(int ii, int jj, int kk, int i0, int j0, int k0) {
    int li = ii - (COMP == 0 ? i0 : i0 - 1);
    int lj = jj - (COMP == 1 ? j0 : j0 - 1);
    int lk = kk - (COMP == 2 ? k0 : k0 - 1);
    const int DI = COMP == 0 ? 9 : 10, DJ = COMP == 1 ? 9 : 10, DK = COMP == 2 ? 9 : 10;
    li = gf::clampi(li, 0, DI - 1);
    lj = gf::clampi(lj, 0, DJ - 1);
    lk = gf::clampi(lk, 0, DK - 1);
    return (li * DJ + lj) * DK + lk;
}

// sample3 of gridmath.py:30-65 on a staged region.
template <int COMP>
__device__ __forceinline__ float sample_region(const float* s, float fx, float fy, float fz,
                                               int nmx, int nmy, int nmz, int i0t, int j0t,
                                               int k0t) {
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
    const float c000 = s[g2p_idx<COMP>(i0, j0, k0, i0t, j0t, k0t)];
    const float c100 = s[g2p_idx<COMP>(i1, j0, k0, i0t, j0t, k0t)];
    const float c010 = s[g2p_idx<COMP>(i0, j1, k0, i0t, j0t, k0t)];
    const float c110 = s[g2p_idx<COMP>(i1, j1, k0, i0t, j0t, k0t)];
    const float c001 = s[g2p_idx<COMP>(i0, j0, k1, i0t, j0t, k0t)];
    const float c101 = s[g2p_idx<COMP>(i1, j0, k1, i0t, j0t, k0t)];
    const float c011 = s[g2p_idx<COMP>(i0, j1, k1, i0t, j0t, k0t)];
    const float c111 = s[g2p_idx<COMP>(i1, j1, k1, i0t, j0t, k0t)];
    const float c00 = c000 * (1.0f - sx) + c100 * sx;
    const float c10 = c010 * (1.0f - sx) + c110 * sx;
    const float c01 = c001 * (1.0f - sx) + c101 * sx;
    const float c11 = c011 * (1.0f - sx) + c111 * sx;
    const float c0 = c00 * (1.0f - sy) + c10 * sy;
    const float c1 = c01 * (1.0f - sy) + c11 * sy;
    return c0 * (1.0f - sz) + c1 * sz;
}

// sample3's interpolation (x, then y, then z) at a corner whose eight values
// sit at base + {0, 1} (k) + {0, SJ} (j) + {0, SI} (i) of a staged region: the
// clamps of sample_region are no-ops for a particle whose corners all lie
// inside the grid and the region (the caller checks).
template <int SJ, int SI>
__device__ __forceinline__ float tri_at(const float* s, int base, float sx, float sy, float sz) {
    const float* b = s + base;
    const float c000 = b[0], c100 = b[SI], c010 = b[SJ], c110 = b[SI + SJ];
    const float c001 = b[1], c101 = b[SI + 1], c011 = b[SJ + 1], c111 = b[SI + SJ + 1];
    const float c00 = c000 * (1.0f - sx) + c100 * sx;
    const float c10 = c010 * (1.0f - sx) + c110 * sx;
    const float c01 = c001 * (1.0f - sx) + c101 * sx;
    const float c11 = c011 * (1.0f - sx) + c111 * sx;
    const float c0 = c00 * (1.0f - sy) + c10 * sy;
    const float c1 = c01 * (1.0f - sy) + c11 * sy;
    return c0 * (1.0f - sz) + c1 * sz;
}

// APIC gather of one component (solver3d.py:1620-1716): clamped corner
// indices, unclamped weights, affine row from world-space face offsets.
template <int COMP>
__device__ __forceinline__ void apic_gather(const float* s, float px, float py, float pz,
                                            float qx, float qy, float qz, int nx, int ny,
                                            int nz, float dx, float inv_dx2_3, int i0t,
                                            int j0t, int k0t, float& vout, float* crow) {
    float fx = qx, fy = qy, fz = qz;
    if (COMP != 0) fx = qx - 0.5f;
    if (COMP != 1) fy = qy - 0.5f;
    if (COMP != 2) fz = qz - 0.5f;
    const int i0 = (int)floorf(fx), j0 = (int)floorf(fy), k0 = (int)floorf(fz);
    const float sx = fx - (float)i0, sy = fy - (float)j0, sz = fz - (float)k0;
    const int imax = COMP == 0 ? nx : nx - 1;
    const int jmax = COMP == 1 ? ny : ny - 1;
    const int kmax = COMP == 2 ? nz : nz - 1;
    float acc = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
#pragma unroll
    for (int di = 0; di < 2; ++di) {
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0) ii = 0;
                if (jj < 0) jj = 0;
                if (kk < 0) kk = 0;
                if (ii > imax) ii = imax;
                if (jj > jmax) jj = jmax;
                if (kk > kmax) kk = kmax;
                const float wxi = di ? sx : 1.0f - sx;
                const float wyi = dj ? sy : 1.0f - sy;
                const float wzi = dk ? sz : 1.0f - sz;
                const float wt = wxi * wyi * wzi;
                const float val = s[g2p_idx<COMP>(ii, jj, kk, i0t, j0t, k0t)];
                acc += wt * val;
                const float ox = (COMP == 0 ? (float)ii : (float)ii + 0.5f) * dx - px;
                const float oy = (COMP == 1 ? (float)jj : (float)jj + 0.5f) * dx - py;
                const float oz = (COMP == 2 ? (float)kk : (float)kk + 0.5f) * dx - pz;
                const float fu = wt * val * inv_dx2_3;
                c0 += fu * ox;
                c1 += fu * oy;
                c2 += fu * oz;
            }
        }
    }
    vout = acc;
    crow[0] = c0;
    crow[1] = c1;
    crow[2] = c2;
}

template <bool APIC>
__global__ __launch_bounds__(GF_WG) void k_g2p_tile(G2PParams a) {
    __shared__ float s_fin[3][G2P_REGION];
    __shared__ float s_sav[APIC ? 1 : 3][APIC ? 1 : G2P_REGION];
    __shared__ float s_p[10][10][10];
    __shared__ int8_t s_mk[10][10][10];
    __shared__ int s_start[64], s_pref[64];
    __shared__ int s_wave[GF_WG / GF_WAVE + 1];

    const int ntiles = a.tnx * a.tny * a.tnz;
    const int t = gf::xcd_remap(blockIdx.x, ntiles);
    const int tk = t % a.tnz, tj = (t / a.tnz) % a.tny, ti = t / (a.tnz * a.tny);
    const int i0 = ti * 8, j0 = tj * 8, k0 = tk * 8;
    const int nx = a.nx, ny = a.ny, nz = a.nz;

    const RowSet rs = build_rows(a.cell_start, ny, nz, i0, min(i0 + 8, nx), j0,
                                 min(j0 + 8, ny), k0, min(k0 + 8, nz), s_start, s_pref,
                                 s_wave);
    // Owned faces can change only next to a fluid cell of this tile or of its
    // lower neighbour tiles.
    bool near_fluid = a.tile_map[t] >= 0 || (ti > 0 && a.tile_map[t - a.tny * a.tnz] >= 0) ||
                      (tj > 0 && a.tile_map[t - a.tnz] >= 0) ||
                      (tk > 0 && a.tile_map[t - 1] >= 0);
    if (rs.total == 0 && !near_fluid) return;

    // The cell box around the tile: markers, then the tiles' slots, then the
    // pressures, each step's loads for all of a thread's cells issued
    // together (out-of-grid or non-fluid cells read an in-bounds stand-in and
    // are masked).
    {
        constexpr int NB = (1000 + GF_WG - 1) / GF_WG;
        int mk[NB], tt[NB], off[NB], sl[NB];
        float pv[NB];
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int i = min((int)threadIdx.x + r * GF_WG, 999);
            const int ca = i / 100, cb = (i / 10) % 10, cc = i % 10;
            const int ci = i0 - 1 + ca, cj = j0 - 1 + cb, ck = k0 - 1 + cc;
            const bool in = ci >= 0 && ci < nx && cj >= 0 && cj < ny && ck >= 0 && ck < nz;
            mk[r] = a.marker[in ? (ci * ny + cj) * nz + ck : 0];
            mk[r] = in ? mk[r] : 0;
            tt[r] = in ? ((ci >> 3) * a.tny + (cj >> 3)) * a.tnz + (ck >> 3) : 0;
            off[r] = ((ci & 7) * 8 + (cj & 7)) * 8 + (ck & 7);
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            sl[r] = a.tile_map[mk[r] == 1 ? tt[r] : 0];
            sl[r] = mk[r] == 1 ? sl[r] : -1;
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const float* src =
                sl[r] >= 0 ? a.pc + (size_t)sl[r] * GF_TILE_CELLS + off[r] : &g_zero_f;
            pv[r] = *src;
        }
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int i = (int)threadIdx.x + r * GF_WG;
            if (i >= 1000) continue;
            const int ca = i / 100, cb = (i / 10) % 10, cc = i % 10;
            s_mk[ca][cb][cc] = (int8_t)mk[r];
            s_p[ca][cb][cc] = pv[r];
        }
    }
    __syncthreads();

    // the tile's scene frame: dx, cell offset and size (one scene: the grid)
    const gf::Frame F = gf::tile_frame(a.tile_scene, a.scene_par, t, a.dx, nx, ny, nz);
    const float scale = a.dt / (a.rho * F.dx);
    const float gdt =
        a.tile_scene ? a.scene_par[gf::SCENE_PAR * max(a.tile_scene[t], 0)] : a.gdt;
    // The face regions: first the saved (and pre-projection) values of all a
    // thread's faces, loads only, so they are in flight together; then the
    // gradient subtract, BC and stores.
    constexpr int NF = (G2P_REGION + GF_WG - 1) / GF_WG;
    float fsav[3][NF], fpre[3][NF];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int DJ = c == 1 ? 9 : 10, DK = c == 2 ? 9 : 10;
        const int oi = c == 0 ? i0 : i0 - 1, oj = c == 1 ? j0 : j0 - 1,
                  ok = c == 2 ? k0 : k0 - 1;
        const int fni = nx + (c == 0), fnj = ny + (c == 1), fnk = nz + (c == 2);
        const float* gsav = c == 0 ? a.us : (c == 1 ? a.vs : a.ws);
        const float* gpre = c == 0 ? a.pu : (c == 1 ? a.pv : a.pw);
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int idx = min((int)threadIdx.x + r * GF_WG, G2P_REGION - 1);
            const int lk = idx % DK, lj = (idx / DK) % DJ, li = idx / (DK * DJ);
            const int ii = oi + li, jj = oj + lj, kk = ok + lk;
            const bool in = ii >= 0 && ii < fni && jj >= 0 && jj < fnj && kk >= 0 && kk < fnk;
            const size_t f = in ? ((size_t)ii * fnj + jj) * fnk + kk : 0;
            fsav[c][r] = gsav[f];
            fpre[c][r] = gpre ? gpre[f] : 0.0f;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int DJ = c == 1 ? 9 : 10, DK = c == 2 ? 9 : 10;
        const int oi = c == 0 ? i0 : i0 - 1, oj = c == 1 ? j0 : j0 - 1,
                  ok = c == 2 ? k0 : k0 - 1;
        const int fni = nx + (c == 0), fnj = ny + (c == 1), fnk = nz + (c == 2);
        float* gfld = c == 0 ? a.u : (c == 1 ? a.v : a.w);
        const float* solid = c == 0 ? a.su : (c == 1 ? a.sv : a.sw);
        const bool has_pre = (c == 0 ? a.pu : (c == 1 ? a.pv : a.pw)) != nullptr;
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int idx = (int)threadIdx.x + r * GF_WG;
            if (idx >= G2P_REGION) continue;
            const int lk = idx % DK, lj = (idx / DK) % DJ, li = idx / (DK * DJ);
            const int ii = oi + li, jj = oj + lj, kk = ok + lk;
            float val = 0.0f, sav = 0.0f;
            if (ii >= 0 && ii < fni && jj >= 0 && jj < fnj && kk >= 0 && kk < fnk) {
                const size_t f = ((size_t)ii * fnj + jj) * fnk + kk;
                // Pre-projection face value: rebuilt from the saved copy exactly
                // as P2G produced it (us [+ g*dt on v], then the BC below), or
                // read from a separate pre-projection array, so the halo never
                // reads u while owners overwrite it.
                sav = fsav[c][r];
                val = has_pre ? fpre[c][r] : (c == 1 ? sav + gdt : sav);
                // Cell coords of the face's two sides in the staged cell region.
                const int ca = ii - (i0 - 1), cb = jj - (j0 - 1), cc = kk - (k0 - 1);
                int la, lb, lc;  // low-side cell
                int n_ax, g_ax;
                if (c == 0) { la = ca - 1; lb = cb; lc = cc; g_ax = ii; n_ax = nx; }
                else if (c == 1) { la = ca; lb = cb - 1; lc = cc; g_ax = jj; n_ax = ny; }
                else { la = ca; lb = cb; lc = cc - 1; g_ax = kk; n_ax = nz; }
                // ca..cc lie in [0,10) for every staged face except the +1
                // planes of the last tile along an axis, which the grid bounds
                // exclude from the gradient (g_ax == n_ax) anyway.
                const bool hi_in = ca < 10 && cb < 10 && cc < 10;
                const bool lo_in = la >= 0 && lb >= 0 && lc >= 0 && la < 10 && lb < 10 && lc < 10;
                const int mlo = lo_in ? s_mk[la][lb][lc] : 0;
                const int mhi = hi_in ? s_mk[ca][cb][cc] : 0;
                const bool lo_side = g_ax == 0 || mlo == 2;
                const bool hi_side = g_ax == n_ax || mhi == 2;
                float sv = 0.0f;
                if (lo_side || hi_side) {
                    sv = solid ? solid[f] : 0.0f;
                    val = sv;
                }
                bool changed = false;
                if (g_ax >= 1 && g_ax <= n_ax - 1) {
                    if ((mlo == 1 || mhi == 1) && mlo != 2 && mhi != 2) {
                        const float plo = mlo == 1 ? s_p[la][lb][lc] : 0.0f;
                        const float phi = mhi == 1 ? s_p[ca][cb][cc] : 0.0f;
                        val = val - scale * (phi - plo);
                        changed = true;
                    }
                }
                if (lo_side || hi_side) val = sv;
                const bool owned = (c == 0 ? (li < 8 || ii == nx) : (li >= 1 && li < 9)) &&
                                   (c == 1 ? (lj < 8 || jj == ny) : (lj >= 1 && lj < 9)) &&
                                   (c == 2 ? (lk < 8 || kk == nz) : (lk >= 1 && lk < 9));
                if (owned && changed) gfld[f] = val;
            }
            s_fin[c][idx] = val;
            if (!APIC) s_sav[c][idx] = sav;
        }
    }
    __syncthreads();

    const float dx = F.dx, dt = a.dt;
    const float eps = dx * 1.001f;
    // samples in the scene's own cell indices: the region origin and the
    // face-array extents of the scene
    const int li0 = i0 - F.ox, lj0 = j0 - F.oy, lk0 = k0 - F.oz;
    const int snx = F.nx, sny = F.ny, snz = F.nz;
    const float inv_dx = 1.0f / dx;
    const float fsnx = (float)(snx - 1), fsny = (float)(sny - 1), fsnz = (float)(snz - 1);
    // clamp box: [1.001 dx, dom - 1.001 dx], or the tile's scene box when packed
    float lox = eps, loy = eps, loz = eps;
    float hix = a.domx - eps, hiy = a.domy - eps, hiz = a.domz - eps;
    if (a.tile_scene) {
        const float* sp = a.scene_par + gf::SCENE_PAR * max(a.tile_scene[t], 0);
        lox = sp[1]; loy = sp[2]; loz = sp[3];
        hix = sp[4]; hiy = sp[5]; hiz = sp[6];
    }
    // The next particle's position and velocity are loaded before this one's
    // are computed and stored (the stores could alias the loads as far as the
    // compiler knows, so otherwise every particle waits a full round trip).
    auto pindex = [&](int q) {
        const int r = row_find(s_pref, rs.nrows, q);
        return s_start[r] + (q - s_pref[r]);
    };
    int nidx = threadIdx.x < rs.total ? pindex(threadIdx.x) : 0;
    float nx_ = 0.0f, ny_ = 0.0f, nz_ = 0.0f, nvx = 0.0f, nvy = 0.0f, nvz = 0.0f;
    if (threadIdx.x < rs.total) {
        nx_ = a.pos[3 * nidx + 0]; ny_ = a.pos[3 * nidx + 1]; nz_ = a.pos[3 * nidx + 2];
        if (!APIC) { nvx = a.vel[3 * nidx + 0]; nvy = a.vel[3 * nidx + 1]; nvz = a.vel[3 * nidx + 2]; }
    }
    for (int q = threadIdx.x; q < rs.total; q += GF_WG) {
        const int pidx = nidx;
        const float px = nx_, py = ny_, pz = nz_;
        const float ovx = nvx, ovy = nvy, ovz = nvz;
        if (q + GF_WG < rs.total) {
            nidx = pindex(q + GF_WG);
            nx_ = a.pos[3 * nidx + 0]; ny_ = a.pos[3 * nidx + 1]; nz_ = a.pos[3 * nidx + 2];
            if (!APIC) { nvx = a.vel[3 * nidx + 0]; nvy = a.vel[3 * nidx + 1]; nvz = a.vel[3 * nidx + 2]; }
        }
        float nu, nv, nw;
        // Fast path (FLIP): a particle at least one cell inside the scene's
        // grid whose corners all fall in the staged regions is sampled without
        // sample3's clamps (they change nothing there), with p/dx as p * (1/dx)
        // and the corner addresses shared by the new and old values of a
        // component.  Other particles take the reference's sample3.
        bool fast = false;
        if (!APIC) {
            const float qx = px * inv_dx, qy = py * inv_dx, qz = pz * inv_dx;
            const float fxo = floorf(qx), fyo = floorf(qy), fzo = floorf(qz);
            const float fxs = floorf(qx - 0.5f), fys = floorf(qy - 0.5f), fzs = floorf(qz - 0.5f);
            // region-local indices: own axis from the tile corner, staggered
            // axes from one below it
            const int ixo = (int)fxo - li0, iyo = (int)fyo - lj0, izo = (int)fzo - lk0;
            const int ixs = (int)fxs - li0 + 1, iys = (int)fys - lj0 + 1, izs = (int)fzs - lk0 + 1;
            fast = qx >= 1.0f && qx <= fsnx && qy >= 1.0f && qy <= fsny && qz >= 1.0f &&
                   qz <= fsnz && (unsigned)ixo <= 7u && (unsigned)iyo <= 7u &&
                   (unsigned)izo <= 7u && (unsigned)ixs <= 8u && (unsigned)iys <= 8u &&
                   (unsigned)izs <= 8u;
            if (fast) {
                const float sxo = qx - fxo, syo = qy - fyo, szo = qz - fzo;
                const float sxs = (qx - 0.5f) - fxs, sys = (qy - 0.5f) - fys,
                            szs = (qz - 0.5f) - fzs;
                const int bu = (ixo * 10 + iys) * 10 + izs;  // u: 9 x 10 x 10
                const int bv = (ixs * 9 + iyo) * 10 + izs;   // v: 10 x 9 x 10
                const int bw = (ixs * 10 + iys) * 9 + izo;   // w: 10 x 10 x 9
                const float pic_u = tri_at<10, 100>(s_fin[0], bu, sxo, sys, szs);
                const float old_u = tri_at<10, 100>(s_sav[0], bu, sxo, sys, szs);
                const float pic_v = tri_at<10, 90>(s_fin[1], bv, sxs, syo, szs);
                const float old_v = tri_at<10, 90>(s_sav[1], bv, sxs, syo, szs);
                const float pic_w = tri_at<9, 90>(s_fin[2], bw, sxs, sys, szo);
                const float old_w = tri_at<9, 90>(s_sav[2], bw, sxs, sys, szo);
                const float fb = a.flip_blend;
                nu = fb * (ovx + (pic_u - old_u)) + (1.0f - fb) * pic_u;
                nv = fb * (ovy + (pic_v - old_v)) + (1.0f - fb) * pic_v;
                nw = fb * (ovz + (pic_w - old_w)) + (1.0f - fb) * pic_w;
            }
        }
        if (!APIC && !fast) {
            const float lx = px / dx, ly = py / dx, lz = pz / dx;
            const float pic_u = sample_region<0>(s_fin[0], lx, ly - 0.5f, lz - 0.5f, snx + 1, sny, snz, li0, lj0, lk0);
            const float old_u = sample_region<0>(s_sav[0], lx, ly - 0.5f, lz - 0.5f, snx + 1, sny, snz, li0, lj0, lk0);
            const float pic_v = sample_region<1>(s_fin[1], lx - 0.5f, ly, lz - 0.5f, snx, sny + 1, snz, li0, lj0, lk0);
            const float old_v = sample_region<1>(s_sav[1], lx - 0.5f, ly, lz - 0.5f, snx, sny + 1, snz, li0, lj0, lk0);
            const float pic_w = sample_region<2>(s_fin[2], lx - 0.5f, ly - 0.5f, lz, snx, sny, snz + 1, li0, lj0, lk0);
            const float old_w = sample_region<2>(s_sav[2], lx - 0.5f, ly - 0.5f, lz, snx, sny, snz + 1, li0, lj0, lk0);
            const float flip_u = ovx + (pic_u - old_u);
            const float flip_v = ovy + (pic_v - old_v);
            const float flip_w = ovz + (pic_w - old_w);
            const float fb = a.flip_blend;
            nu = fb * flip_u + (1.0f - fb) * pic_u;
            nv = fb * flip_v + (1.0f - fb) * pic_v;
            nw = fb * flip_w + (1.0f - fb) * pic_w;
        } else if (APIC) {
            const float lx = px / dx, ly = py / dx, lz = pz / dx;
            const float inv_dx2_3 = 3.0f / (dx * dx);
            float cm[9];
            apic_gather<0>(s_fin[0], px, py, pz, lx, ly, lz, snx, sny, snz, dx, inv_dx2_3, li0, lj0, lk0, nu, cm + 0);
            apic_gather<1>(s_fin[1], px, py, pz, lx, ly, lz, snx, sny, snz, dx, inv_dx2_3, li0, lj0, lk0, nv, cm + 3);
            apic_gather<2>(s_fin[2], px, py, pz, lx, ly, lz, snx, sny, snz, dx, inv_dx2_3, li0, lj0, lk0, nw, cm + 6);
#pragma unroll
            for (int c = 0; c < 9; ++c) a.C[9 * pidx + c] = cm[c];
        }
        a.vel[3 * pidx + 0] = nu;
        a.vel[3 * pidx + 1] = nv;
        a.vel[3 * pidx + 2] = nw;
        float npx = px + nu * dt, npy = py + nv * dt, npz = pz + nw * dt;
        npx = fminf(fmaxf(npx, lox), hix);
        npy = fminf(fmaxf(npy, loy), hiy);
        npz = fminf(fmaxf(npz, loz), hiz);
        a.pos[3 * pidx + 0] = npx;
        a.pos[3 * pidx + 1] = npy;
        a.pos[3 * pidx + 2] = npz;
    }
}

__global__ __launch_bounds__(GF_WG) void k_compact_to_dense(const float* __restrict__ src,
                                                            const int32_t* __restrict__ tile_map,
                                                            float* __restrict__ dst, int nx,
                                                            int ny, int nz, int tny, int tnz) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny * nz) return;
    const int k = c % nz, j = (c / nz) % ny, i = c / (nz * ny);
    const int slot = tile_map[((i >> 3) * tny + (j >> 3)) * tnz + (k >> 3)];
    dst[c] = slot >= 0 ? src[(size_t)slot * GF_TILE_CELLS + (((i & 7) * 8 + (j & 7)) * 8 + (k & 7))]
                       : 0.0f;
}

__global__ __launch_bounds__(GF_WG) void k_unsort(const float* __restrict__ in,
                                                  const int32_t* __restrict__ pid, int n,
                                                  int width, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = pid[i];
    for (int c = 0; c < width; ++c) out[(size_t)d * width + c] = in[(size_t)i * width + c];
}

// Divergence (solver3d.py:846-871) of the dense u,v,w into the compact rhs of
// the active tiles, the same arithmetic as the fused P2G's; for stages that
// change the velocities between P2G and the pressure solve.  One workgroup
// per tile, two cells per thread.
__global__ __launch_bounds__(GF_WG) void k_scene_divergence(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
    const int32_t* __restrict__ marker, const int32_t* __restrict__ tile_map,
    float* __restrict__ b, int nx, int ny, int nz, int tny, int tnz, float dx, float dt,
    float rho) {
    const int t = blockIdx.x;
    const int slot = tile_map[t];
    if (slot < 0) return;
    const int tk = t % tnz, tj = (t / tnz) % tny, ti = t / (tnz * tny);
    const float scale = rho * dx / dt;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int l = threadIdx.x + h * GF_WG;
        const int i = ti * 8 + (l >> 6), j = tj * 8 + ((l >> 3) & 7), k = tk * 8 + (l & 7);
        float dv = 0.0f;
        if (i < nx && j < ny && k < nz) {
            const size_t c = ((size_t)i * ny + j) * nz + k;
            if (marker[c] == 1) {
                const float u0 = u[c], u1 = u[c + (size_t)ny * nz];
                const size_t cv = ((size_t)i * (ny + 1) + j) * nz + k;
                const float v0 = v[cv], v1 = v[cv + nz];
                const size_t cw = ((size_t)i * ny + j) * (nz + 1) + k;
                const float w0 = w[cw], w1 = w[cw + 1];
                const float d = u1 - u0 + v1 - v0 + w1 - w0;
                dv = d * scale;
            }
        }
        b[(size_t)slot * GF_TILE_CELLS + l] = dv;
    }
}

G2PParams g2p_params(const gf_scene_t* s, const float* pressure, float dt, float rho,
                     float flip_blend, float gdt) {
    G2PParams a;
    a.pos = s->pos; a.vel = s->vel; a.C = s->affine_c;
    a.cell_start = s->cell_start; a.marker = s->marker; a.tile_map = s->tile_map;
    a.u = s->u; a.v = s->v; a.w = s->w; a.us = s->us; a.vs = s->vs; a.ws = s->ws;
    a.su = s->solid_u; a.sv = s->solid_v; a.sw = s->solid_w;
    a.pc = pressure;
    a.pu = nullptr; a.pv = nullptr; a.pw = nullptr;
    a.nx = s->nx; a.ny = s->ny; a.nz = s->nz; a.tnx = s->tnx; a.tny = s->tny; a.tnz = s->tnz;
    a.dx = s->dx; a.dt = dt; a.rho = rho; a.flip_blend = flip_blend; a.gdt = gdt;
    a.domx = s->dom[0]; a.domy = s->dom[1]; a.domz = s->dom[2];
    const bool packed = s->n_scenes > 1;
    a.tile_scene = packed ? s->tile_scene : nullptr;
    a.scene_par = packed ? s->scene_par : nullptr;
    return a;
}

int g2p_check(const gf_scene_t* s, const float* pressure, int transfer_mode, float rho,
              const char* who) {
    GF_REQUIRE(transfer_mode == GF_TRANSFER_FLIP || transfer_mode == GF_TRANSFER_APIC,
               "%s: transfer_mode %d", who, transfer_mode);
    GF_REQUIRE(s->u && s->v && s->w && s->marker && s->cell_start && s->tile_map && pressure,
               "%s: grid buffers not set", who);
    GF_REQUIRE(rho != 0.0f, "%s: rho must be nonzero", who);
    GF_REQUIRE(s->us && s->vs && s->ws, "%s: us/vs/ws not set", who);
    if (s->n > 0) GF_REQUIRE(s->pos && s->vel, "%s: particles not set", who);
    if (transfer_mode == GF_TRANSFER_APIC && s->n > 0)
        GF_REQUIRE(s->affine_c, "%s: APIC needs affine_c", who);
    return 0;
}

int g2p_launch(const gf_scene_t* s, const G2PParams& a, int transfer_mode, void* stream) {
    const int ntiles = s->tnx * s->tny * s->tnz;
    hipStream_t st = (hipStream_t)stream;
    if (transfer_mode == GF_TRANSFER_APIC)
        hipLaunchKernelGGL(k_g2p_tile<true>, dim3(ntiles), dim3(GF_WG), 0, st, a);
    else
        hipLaunchKernelGGL(k_g2p_tile<false>, dim3(ntiles), dim3(GF_WG), 0, st, a);
    GF_LAUNCH_CHECK("k_g2p_tile");
    return 0;
}

P2GParams p2g_params(const gf_scene_t* s, float g, float dt, float rho) {
    P2GParams a;
    a.pos = s->pos; a.vel = s->vel; a.C = s->affine_c;
    a.cell_start = s->cell_start; a.marker = s->marker; a.tile_map = s->tile_map;
    a.u = s->u; a.v = s->v; a.w = s->w; a.us = s->us; a.vs = s->vs; a.ws = s->ws;
    a.uw = nullptr; a.vw = nullptr; a.ww = nullptr;
    a.su = s->solid_u; a.sv = s->solid_v; a.sw = s->solid_w;
    a.b = s->b;
    const bool packed = s->n_scenes > 1;
    a.tile_scene = packed ? s->tile_scene : nullptr;
    a.scene_par = packed ? s->scene_par : nullptr;
    a.tile_occ = s->tile_occ;
    a.flags = s->flags;
    a.nx = s->nx; a.ny = s->ny; a.nz = s->nz; a.tnx = s->tnx; a.tny = s->tny; a.tnz = s->tnz;
    a.dx = s->dx; a.g = g; a.dt = dt; a.rho = rho;
    return a;
}

}  // namespace

namespace gf {
int scene_check(const gf_scene_t* s, const char* who);
}

extern "C" int gf_scene_p2g(const gf_scene_t* s, int transfer_mode, float gravity, float dt,
                            float rho, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_p2g");
    if (rc) return rc;
    GF_REQUIRE(s->n_scenes <= 1, "gf_scene_p2g: not for packed scenes (their particles are in "
                                 "per-scene units; use the staged P2G and gf_scene_g2p)");
    GF_REQUIRE(transfer_mode == GF_TRANSFER_FLIP || transfer_mode == GF_TRANSFER_APIC,
               "gf_scene_p2g: transfer_mode %d", transfer_mode);
    GF_REQUIRE(s->u && s->v && s->w && s->us && s->vs && s->ws && s->b && s->marker &&
                   s->cell_start && s->tile_map,
               "gf_scene_p2g: grid buffers not set");
    GF_REQUIRE(dt != 0.0f, "gf_scene_p2g: dt must be nonzero");
    if (s->n > 0) GF_REQUIRE(s->pos && s->vel, "gf_scene_p2g: particles not set");
    if (transfer_mode == GF_TRANSFER_APIC && s->n > 0)
        GF_REQUIRE(s->affine_c, "gf_scene_p2g: APIC needs affine_c");
    P2GParams a = p2g_params(s, gravity, dt, rho);
    const int ntiles = s->tnx * s->tny * s->tnz;
    hipStream_t st = (hipStream_t)stream;
    if (transfer_mode == GF_TRANSFER_APIC)
        hipLaunchKernelGGL((k_p2g_tile<true, false>), dim3(ntiles), dim3(GF_WG), 0, st, a);
    else
        hipLaunchKernelGGL((k_p2g_tile<false, false>), dim3(ntiles), dim3(GF_WG), 0, st, a);
    GF_LAUNCH_CHECK("k_p2g_tile");
    return 0;
}

extern "C" int gf_scene_p2g_stage_size(int nx, int ny, int nz, int64_t* stage_floats,
                                       int64_t* flag_ints) {
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && stage_floats && flag_ints,
               "gf_scene_p2g_stage_size: bad arguments");
    const int64_t ntiles = (int64_t)((nx + 7) / 8) * ((ny + 7) / 8) * ((nz + 7) / 8);
    *stage_floats = ntiles * STAGE_PER_TILE;
    *flag_ints = ntiles;
    return 0;
}

extern "C" int gf_scene_p2g_staged(const gf_scene_t* s, int transfer_mode, float gravity,
                                   float dt, float rho, float* stage, int32_t* has,
                                   void* stream) {
    int rc = gf::scene_check(s, "gf_scene_p2g_staged");
    if (rc) return rc;
    GF_REQUIRE(transfer_mode == GF_TRANSFER_FLIP || transfer_mode == GF_TRANSFER_APIC,
               "gf_scene_p2g_staged: transfer_mode %d", transfer_mode);
    GF_REQUIRE(s->u && s->v && s->w && s->us && s->vs && s->ws && s->b && s->marker &&
                   s->cell_start && s->tile_map && s->tile_occ && s->flags && stage && has,
               "gf_scene_p2g_staged: buffers not set");
    GF_REQUIRE(dt != 0.0f, "gf_scene_p2g_staged: dt must be nonzero");
    if (s->n > 0) GF_REQUIRE(s->pos && s->vel, "gf_scene_p2g_staged: particles not set");
    if (transfer_mode == GF_TRANSFER_APIC && s->n > 0)
        GF_REQUIRE(s->affine_c, "gf_scene_p2g_staged: APIC needs affine_c");
    P2GParams a = p2g_params(s, gravity, dt, rho);
    const int ntiles = s->tnx * s->tny * s->tnz;
    hipStream_t st = (hipStream_t)stream;
    const bool apic = transfer_mode == GF_TRANSFER_APIC, packed = a.tile_scene != nullptr;
    auto own = apic ? (packed ? k_p2g_cells_apic<true> : k_p2g_cells_apic<false>)
                    : (packed ? k_p2g_cells_flip<true> : k_p2g_cells_flip<false>);
    hipLaunchKernelGGL(own, dim3(ntiles), dim3(64 * P2GC_NW), 0, st, a, stage, has);
    hipLaunchKernelGGL(k_p2g_finish, dim3(ntiles), dim3(GF_WG), 0, st, a, stage, has);
    GF_LAUNCH_CHECK("k_p2g_cells/k_p2g_finish");
    return 0;
}

extern "C" int gf_scene_p2g_raw(const gf_scene_t* s, int transfer_mode, float* uw, float* vw,
                                float* ww, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_p2g_raw");
    if (rc) return rc;
    GF_REQUIRE(s->n_scenes <= 1, "gf_scene_p2g_raw: not for packed scenes (their particles are in "
                                 "per-scene units; use the staged P2G and gf_scene_g2p)");
    GF_REQUIRE(transfer_mode == GF_TRANSFER_FLIP || transfer_mode == GF_TRANSFER_APIC,
               "gf_scene_p2g_raw: transfer_mode %d", transfer_mode);
    GF_REQUIRE(s->u && s->v && s->w && uw && vw && ww && s->cell_start,
               "gf_scene_p2g_raw: grid buffers not set");
    if (s->n > 0) GF_REQUIRE(s->pos && s->vel, "gf_scene_p2g_raw: particles not set");
    if (transfer_mode == GF_TRANSFER_APIC && s->n > 0)
        GF_REQUIRE(s->affine_c, "gf_scene_p2g_raw: APIC needs affine_c");
    P2GParams a = p2g_params(s, 0.0f, 1.0f, 1.0f);
    a.uw = uw; a.vw = vw; a.ww = ww;
    const int ntiles = s->tnx * s->tny * s->tnz;
    hipStream_t st = (hipStream_t)stream;
    if (transfer_mode == GF_TRANSFER_APIC)
        hipLaunchKernelGGL((k_p2g_tile<true, true>), dim3(ntiles), dim3(GF_WG), 0, st, a);
    else
        hipLaunchKernelGGL((k_p2g_tile<false, true>), dim3(ntiles), dim3(GF_WG), 0, st, a);
    GF_LAUNCH_CHECK("k_p2g_tile(raw)");
    return 0;
}

extern "C" int gf_scene_g2p(const gf_scene_t* s, const float* pressure, int transfer_mode,
                            float gravity, float dt, float rho, float flip_blend, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_g2p");
    if (rc) return rc;
    rc = g2p_check(s, pressure, transfer_mode, rho, "gf_scene_g2p");
    if (rc) return rc;
    return g2p_launch(s, g2p_params(s, pressure, dt, rho, flip_blend, gravity * dt),
                      transfer_mode, stream);
}

extern "C" int gf_scene_g2p_pre(const gf_scene_t* s, const float* pressure, const float* pre_u,
                                const float* pre_v, const float* pre_w, int transfer_mode,
                                float dt, float rho, float flip_blend, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_g2p_pre");
    if (rc) return rc;
    GF_REQUIRE(s->n_scenes <= 1, "gf_scene_g2p_pre: not for packed scenes (their particles are in "
                                 "per-scene units; use the staged P2G and gf_scene_g2p)");
    rc = g2p_check(s, pressure, transfer_mode, rho, "gf_scene_g2p_pre");
    if (rc) return rc;
    GF_REQUIRE(pre_u && pre_v && pre_w, "gf_scene_g2p_pre: pre-projection faces not set");
    GF_REQUIRE(pre_u != s->u && pre_v != s->v && pre_w != s->w,
               "gf_scene_g2p_pre: pre-projection faces must not alias u, v, w");
    G2PParams a = g2p_params(s, pressure, dt, rho, flip_blend, 0.0f);
    a.pu = pre_u; a.pv = pre_v; a.pw = pre_w;
    return g2p_launch(s, a, transfer_mode, stream);
}

extern "C" int gf_scene_divergence(const gf_scene_t* s, float dt, float rho, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_divergence");
    if (rc) return rc;
    GF_REQUIRE(s->n_scenes <= 1, "gf_scene_divergence: not for packed scenes (their particles are in "
                                 "per-scene units; use the staged P2G and gf_scene_g2p)");
    GF_REQUIRE(s->u && s->v && s->w && s->marker && s->tile_map && s->b,
               "gf_scene_divergence: grid buffers not set");
    GF_REQUIRE(dt != 0.0f, "gf_scene_divergence: dt must be nonzero");
    const int ntiles = s->tnx * s->tny * s->tnz;
    hipLaunchKernelGGL(k_scene_divergence, dim3(ntiles), dim3(GF_WG), 0, (hipStream_t)stream,
                       s->u, s->v, s->w, s->marker, s->tile_map, s->b, s->nx, s->ny, s->nz,
                       s->tny, s->tnz, s->dx, dt, rho);
    GF_LAUNCH_CHECK("k_scene_divergence");
    return 0;
}

extern "C" int gf_scene_compact_to_dense(const gf_scene_t* s, const float* compact,
                                         float* dense, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_compact_to_dense");
    if (rc) return rc;
    GF_REQUIRE(compact && dense && s->tile_map, "gf_scene_compact_to_dense: null buffer");
    const int C = s->nx * s->ny * s->nz;
    hipLaunchKernelGGL(k_compact_to_dense, dim3((C + GF_WG - 1) / GF_WG), dim3(GF_WG), 0,
                       (hipStream_t)stream, compact, s->tile_map, dense, s->nx, s->ny, s->nz,
                       s->tny, s->tnz);
    GF_LAUNCH_CHECK("k_compact_to_dense");
    return 0;
}

extern "C" int gf_scene_unsort(const gf_scene_t* s, const float* in, int width, float* out,
                               void* stream) {
    int rc = gf::scene_check(s, "gf_scene_unsort");
    if (rc) return rc;
    GF_REQUIRE(width >= 1 && width <= 16, "gf_scene_unsort: width %d", width);
    if (s->n == 0) return 0;
    GF_REQUIRE(in && out && s->pid, "gf_scene_unsort: null buffer");
    hipLaunchKernelGGL(k_unsort, dim3((s->n + GF_WG - 1) / GF_WG), dim3(GF_WG), 0,
                       (hipStream_t)stream, in, s->pid, s->n, width, out);
    GF_LAUNCH_CHECK("k_unsort");
    return 0;
}
