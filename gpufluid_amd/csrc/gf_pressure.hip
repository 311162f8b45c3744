// Pressure projection on the compacted active 8^3 tiles.
//
// Storage: every vector holds 512 entries per active tile (slot-major, cell
// index (li*8+lj)*8+lk), so a wavefront streams one 64-cell li-plane.  Entries
// of non-fluid cells are kept at zero; the 7-point operator
// A x = diag*x - sum_{fluid nb} x of k3_apply_A (solver3d.py:1743-1797) then
// equals diag*x minus the sum over all six neighbours, with diag = number of
// in-grid non-solid neighbours stored in the flags byte.
//
// PCG (solver3d.py:2917-3043 restated): Jacobi preconditioner M = max(diag,1)
// (k3_compute_diag, 1803-1823), x0 = 0, stop on ||r||_2 < tol*||b||_2,
// alpha = 0 when |pAp| < 1e-30 and beta = 0 when |rz| < 1e-30 (1883-1896).
// A kernel boundary is the only grid-wide sync here, and two forms of the
// iteration are built:
//  * two kernels per iteration (standard PCG, the default above
//    PCG_CG1_MAX_TILES active tiles): a direction+SpMV kernel
//    (p = z + beta p, Ap, p.Ap) and an update kernel (x += alpha p,
//    r -= alpha Ap, z = M^-1 r, r.z, r.r); r is held as z = M^-1 r;
//  * one kernel per iteration (Chronopoulos-Gear form, below that size, where
//    the fixed cost of a launch dominates), which needs one global reduction
//    per iteration instead of two:
//   gamma = (r,u), delta = (Au,u)        (reduced by the previous kernel)
//   beta = gamma/gamma_old, alpha = gamma / (delta - beta*gamma/alpha_old)
//   p = u + beta p,  s~ = w~ + beta s~,  x += alpha p,  u -= alpha s~,
//   w~ = M^-1 A u  (stencil: the neighbours' new u is recomputed from their
//                   u, w~, s~, ping-ponged),  partials of the next gamma,
//   delta and r.r
// with u = M^-1 r, w~ = M^-1 A u, s~ = M^-1 A p stored (r = M u is never
// stored).  In exact arithmetic the iterates equal standard PCG's.  Dot
// products are FP64 with a deterministic two-level reduction.  In the
// mixed-precision mode the vectors are FP32 and an FP64 iterate x64 + x32 is
// kept; when ||r|| has dropped about 32x (r.r 1000x) since the last
// replacement, and before convergence is accepted, the residual vectors (z;
// or u, s~ and w~) are rebuilt from the FP64 true residual b - A(x64+x32)
// (reliable updates, Sleijpen & van der Vorst 1996).
#include "gf_common.h"
#include <hip/hip_ext.h>
#include <atomic>
#include <string.h>
#include <mutex>
#include <type_traits>

namespace {

// Solver scalars, one log per scene.  Kernels form a fixed launch sequence
// numbered L; launch L reads entry (L-1)&1 of its scene's pair and writes
// entry L&1 (the scene's first workgroup only), so no kernel reads a field
// that it also writes.  Entries 2S.. hold each scene's final log.
struct PcgLog {
    double rz;       // r.z of the current residual (used by the next alpha)
    double rz_old;   // r.z used by the last alpha (denominator of the next beta)
    double rr, bb, alpha, alpha2, rr_upd, rr_true, pap;
    int iters, done, converged, maxed, upd, merge, nupd, pending;
    int fresh;       // 1: the update kernel's r.z / r.r partials are not consumed yet
    int pbuf;        // direction buffer holding the current p
    int xpend;       // two-kernel form: x lacks alpha p of the last xpend (<= 3) updates,
                     // whose alphas are alpha, alpha2, alpha3 (latest first)
    int pad_i;
    double alpha3;
};
static_assert(sizeof(PcgLog) == 16 * sizeof(double), "log entry is 16 doubles");

// Per-workgroup partial sums, PCG_GMAX entries per kind, in two sets chosen by
// the parity of the writing launch: a kernel may read a kind that it also
// writes (the iteration kernel consumes and produces gamma, delta, r.r), and
// without the second set a workgroup that finishes early would overwrite its
// partial before a late-starting workgroup of the same launch had read it.
enum { P_PAP = 0, P_RZ, P_RR, P_BB, P_RRT, P_RZT, P_KINDS };
constexpr int P_SET = 8;  // kinds per parity set (P_KINDS rounded up)
constexpr int PCG_GMAX = 2048;
enum { PCG_STD = 1, PCG_CG1 = 2 };
// Below this many active tiles the one-kernel iteration is faster (its fixed
// per-launch cost is paid once per iteration); above it the two-kernel form
// streams more efficiently (measured on MI355X, us per iteration, one vs two
// kernels: 224 tiles 10.6 vs 15.8, 1.7 K tiles 14.5 vs 17.0, 3.2 K tiles 20.8
// vs 21.2, 5.8 K tiles 29.9 vs 28.1, 13.3 K tiles 65 vs 55).
constexpr int PCG_CG1_MAX_TILES = 3072;
constexpr int PCG_DEFAULT_G = 1024;

struct PcgArgs {
    const uint8_t* flags;
    const int32_t* nbr;
    const float* b;
    float* u32[2]; float* w32[2]; float* s32[2];   // ping-pong u, w~, s~
    float* p32; float* x32;
    double* u64[2]; double* w64[2]; double* s64[2];
    double* p64; double* x64;
    float* pc;
    // two-kernel form: the cells' flags at 4 bits per cell (one 16-bit word per
    // quad, 128 per tile), written by its setup into the buffer the one-kernel
    // form uses for w~ (unused by this form); read by every iteration
    uint16_t* pflags;
    double* part;        // [P_KINDS][PCG_GMAX]
    PcgLog* log;         // [2*S] live pairs
    PcgLog* res;         // [S] results (gf_scene_pcg_stats reads them)
    PcgLog* hout;        // when set: the log pairs are also written here (pinned host memory)
    int32_t* part_info;  // [8*PCG_GMAX] per workgroup: scene, wb, we, t0, t1; then [S+1] wg offsets
    const int32_t* scene_slot;  // [S+1] compact tile offsets per scene (NULL: one scene, nact tiles)
    const int32_t* nact_dev;    // one scene, count on the device ([0] active tiles) or NULL
    int nact;            // active tiles (one scene, nact_dev NULL)
    int S;               // scenes
    int variant;         // PCG_STD (two kernels per iteration) or PCG_CG1 (one)
    int G;               // workgroups; a scene's workgroups split its tiles evenly
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

// Outputs use plain stores.  Non-temporal stores measured the same for the
// two-kernel PCG and slower for the one-kernel form at 3-6 K tiles (31 vs 39
// us per iteration at 5.8 K tiles), whose vectors are re-read from the
// last-level cache by the next launch.

__device__ __forceinline__ uint32_t ld_flags(const uint8_t* f, size_t idx) {
    return *reinterpret_cast<const uint32_t*>(f + idx);
}
__device__ __forceinline__ int fl_fluid(uint32_t fw, int e) { return (fw >> (8 * e + 3)) & 1; }

// A quad's four flag bytes (each < 16: fluid bit 3, diag bits 0-2) as one
// 16-bit word of nibbles, and back.
__device__ __forceinline__ uint32_t pack_flags(uint32_t fw) {
    return (fw & 0xFu) | ((fw >> 4) & 0xF0u) | ((fw >> 8) & 0xF00u) | ((fw >> 12) & 0xF000u);
}
__device__ __forceinline__ uint32_t unpack_flags(uint32_t n) {
    return (n & 0xFu) | ((n & 0xF0u) << 4) | ((n & 0xF00u) << 8) | ((n & 0xF000u) << 12);
}

// v if keep, else +0, as a bit mask: a select here would let the compiler
// sink the loads that feed v into a branch and serialise them.
__device__ __forceinline__ float keep_if(float v, int keep) {
    return __uint_as_float(__float_as_uint(v) & (0u - (uint32_t)keep));
}
__device__ __forceinline__ double keep_if(double v, int keep) {
    return __longlong_as_double(__double_as_longlong(v) & (0ll - (long long)keep));
}
__device__ __forceinline__ int fl_diag(uint32_t fw, int e) { return (fw >> (8 * e)) & 7; }

// Element offsets of a quad's neighbours: the four -x/+x/-y/+y quads and the
// two z cells outside the quad (-z of its first cell, +z of its last).  A
// neighbour in an inactive tile gets the quad's own offset and a cleared ok
// bit, so every later load is unconditional and in bounds.  The tile's
// neighbour table is read unconditionally (t is wave-uniform) and the
// addresses are formed with selects: no load sits behind a branch.
struct NbIdx {
    size_t xm, xp, ym, yp, zm, zp;
    uint32_t ok;  // bit d set: neighbour d (order -x,+x,-y,+y,-z,+z) exists
};

typedef const __attribute__((address_space(4))) int32_t gf_const_i32;

__device__ __forceinline__ NbIdx nb_index(const int32_t* __restrict__ nbr, int t, size_t base,
                                          const Quad& Q) {
    gf_const_i32* row = (gf_const_i32*)(nbr + (size_t)t * 6);
    const int2 n01 = make_int2(row[0], row[1]), n23 = make_int2(row[2], row[3]),
               n45 = make_int2(row[4], row[5]);
    const size_t own = base + Q.l0;
    NbIdx o;
    const bool bxm = Q.li > 0, bxp = Q.li < 7, bym = Q.lj > 0, byp = Q.lj < 7;
    const bool bzm = Q.kh == 1, bzp = Q.kh == 0;
    const bool kxm = bxm || n01.x >= 0, kxp = bxp || n01.y >= 0;
    const bool kym = bym || n23.x >= 0, kyp = byp || n23.y >= 0;
    const bool kzm = bzm || n45.x >= 0, kzp = bzp || n45.y >= 0;
    o.xm = bxm ? own - 64 : (kxm ? (size_t)n01.x * 512 + Q.l0 + 448 : own);
    o.xp = bxp ? own + 64 : (kxp ? (size_t)n01.y * 512 + Q.l0 - 448 : own);
    o.ym = bym ? own - 8 : (kym ? (size_t)n23.x * 512 + Q.l0 + 56 : own);
    o.yp = byp ? own + 8 : (kyp ? (size_t)n23.y * 512 + Q.l0 - 56 : own);
    o.zm = bzm ? own - 1 : (kzm ? (size_t)n45.x * 512 + Q.l0 + 7 : own);
    o.zp = bzp ? own + 4 : (kzp ? (size_t)n45.y * 512 + Q.l0 - 4 : own);
    o.ok = (uint32_t)kxm | (uint32_t)kxp << 1 | (uint32_t)kym << 2 | (uint32_t)kyp << 3 |
           (uint32_t)kzm << 4 | (uint32_t)kzp << 5;
    return o;
}

// Neighbour values of a quad (see NbIdx); neighbours in inactive tiles read as 0.
template <typename T>
struct Nb4 {
    T xm[4], xp[4], ym[4], yp[4];
    T zm0, zp3;
};

template <typename T>
__device__ __forceinline__ void load_nb4(const T* __restrict__ v, const NbIdx& n, Nb4<T>& o) {
    ld4(v + n.xm, o.xm);
    ld4(v + n.xp, o.xp);
    ld4(v + n.ym, o.ym);
    ld4(v + n.yp, o.yp);
    o.zm0 = v[n.zm];
    o.zp3 = v[n.zp];
    for (int e = 0; e < 4; ++e) {
        o.xm[e] = (n.ok & 1) ? o.xm[e] : (T)0;
        o.xp[e] = (n.ok & 2) ? o.xp[e] : (T)0;
        o.ym[e] = (n.ok & 4) ? o.ym[e] : (T)0;
        o.yp[e] = (n.ok & 8) ? o.yp[e] : (T)0;
    }
    o.zm0 = (n.ok & 16) ? o.zm0 : (T)0;
    o.zp3 = (n.ok & 32) ? o.zp3 : (T)0;
}

// ---- neighbours through lane exchange -----------------------------------
// A wave holds four li-planes of one tile (lane = 16*(li&3) + 2*lj + kh), so
// most stencil neighbours of a quad sit in another lane of the same wave:
// -y/+y two lanes away in the same 16-lane row (DPP row shifts), -z/+z in the
// partner lane (DPP quad permute), -x/+x sixteen lanes away (ds_bpermute).
// Only the values across a face of the wave's block are loaded: the -x plane
// below the wave's first plane (previous tile or the other wave's half), the
// +x plane above its last, the -y row / +y row of the neighbouring tile, and
// the -z / +z cell of the neighbouring tile.  For each direction the tile
// that holds the face values is the same for the whole wave and the lane's
// offset inside it does not change from tile to tile, so a face load is a
// scalar base (per tile) plus a 32-bit lane offset fixed for the launch: no
// per-lane address arithmetic in the loop.  An absent neighbour reads a tile
// of zeros.  The kernels are VALU-bound at these sizes (measured: ~1500 VALU
// instructions per wave against ~45% of wave time waiting on memory), which
// is what this layout is for.
struct FaceOff {
    uint32_t own;                     // the quad's own offset (Q.l0)
    uint32_t xm, xp, ym, yp, zm, zp;  // element offsets inside the holding tile
};

__device__ __forceinline__ FaceOff face_off(const Quad& Q) {
    const int half = Q.li >> 2;  // wave holds planes 4*half .. 4*half+3
    const uint32_t pos = Q.lj * 8 + Q.kh * 4;
    FaceOff o;
    o.own = Q.l0;
    // -x of plane 4*half: plane 3 of this tile (half 1) or plane 7 of the -x tile
    o.xm = (half ? 3 * 64 : 7 * 64) + pos;
    // +x of plane 4*half+3: plane 4 of this tile (half 0) or plane 0 of the +x tile
    o.xp = (half ? 0 : 4 * 64) + pos;
    o.ym = (Q.li * 8 + 7) * 8 + Q.kh * 4;
    o.yp = (Q.li * 8) * 8 + Q.kh * 4;
    o.zm = (Q.li * 8 + Q.lj) * 8 + 7;
    o.zp = (Q.li * 8 + Q.lj) * 8;
    return o;
}

struct FaceTiles {
    int xm, xp, ym, yp, zm, zp;  // holding tile per direction (wave-uniform; -1: absent)
};

// The table row is read through the constant address space: t is
// wave-uniform and the table is never written by these kernels, so the row
// comes in with scalar loads (measured 1-4 us less per launch at 13 K tiles
// than per-lane loads).
__device__ __forceinline__ FaceTiles face_tiles(const int32_t* __restrict__ nbr, int t) {
    gf_const_i32* row = (gf_const_i32*)(nbr + (size_t)t * 6);
    const int half = (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6) & 1;
    // the whole row is read (one scalar load, no branch), then selected
    const int r0 = row[0], r1 = row[1];
    FaceTiles f;
    f.xm = half ? t : r0;
    f.xp = half ? r1 : t;
    f.ym = row[2];
    f.yp = row[3];
    f.zm = row[4];
    f.zp = row[5];
    return f;
}

// Stands in for the vector entries of an absent neighbour tile (read as FP32
// or FP64; never written).
__device__ double g_zero_tile[GF_TILE_CELLS] = {};

template <typename T>
__device__ __forceinline__ const T* tile_base(const T* v, int tile) {
    return tile >= 0 ? v + (size_t)tile * GF_TILE_CELLS : reinterpret_cast<const T*>(g_zero_tile);
}

// The packed flags of the quad at element offset o of tile t (t < 0: none).
__device__ __forceinline__ uint32_t ld_pflags(const uint16_t* pf, int t, uint32_t o) {
    const uint16_t* base = t >= 0 ? pf + (size_t)t * (GF_TILE_CELLS / 4)
                                  : reinterpret_cast<const uint16_t*>(g_zero_tile);
    return unpack_flags(base[o >> 2]);
}

// Neighbour-table row of no tile (every neighbour absent): the first-pass
// prefetch of a workgroup with no tiles reads the zero tile through it.
__device__ int32_t g_no_nbr[8] = {-1, -1, -1, -1, -1, -1, -1, -1};

// The neighbour row of tile t as stored (t < 0: every neighbour absent); the
// wave's own -x or +x entry is put in by for_half when the row is used, so
// a row loaded one pass ahead is not waited for until then.
__device__ __forceinline__ FaceTiles face_row(const int32_t* __restrict__ nbr, int t) {
    gf_const_i32* row = (gf_const_i32*)(t >= 0 ? nbr + (size_t)t * 6 : g_no_nbr);
    FaceTiles f;
    f.xm = row[0];
    f.xp = row[1];
    f.ym = row[2];
    f.yp = row[3];
    f.zm = row[4];
    f.zp = row[5];
    return f;
}

// face_tiles of tile t from its stored row.
__device__ __forceinline__ FaceTiles for_half(FaceTiles f, int t) {
    const int half = (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6) & 1;
    f.xm = half ? t : f.xm;
    f.xp = half ? f.xp : t;
    return f;
}

// Face values of one vector.
template <typename T>
__device__ __forceinline__ void load_face(const T* __restrict__ v, const FaceTiles& f,
                                          const FaceOff& off, Nb4<T>& o) {
    ld4(tile_base(v, f.xm) + off.xm, o.xm);
    ld4(tile_base(v, f.xp) + off.xp, o.xp);
    ld4(tile_base(v, f.ym) + off.ym, o.ym);
    ld4(tile_base(v, f.yp) + off.yp, o.yp);
    o.zm0 = tile_base(v, f.zm)[off.zm];
    o.zp3 = tile_base(v, f.zp)[off.zp];
}

// Lanes whose DPP source lies outside the row read 0 (bound_ctrl); they take
// a face value instead.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ float lane_read(float v, int byte_addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(byte_addr, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int byte_addr) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(byte_addr, (int)b);
    const int hi = __builtin_amdgcn_ds_bpermute(byte_addr, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (uint32_t)lo);
}

enum { DPP_ROW_SHL2 = 0x102, DPP_ROW_SHR2 = 0x112, DPP_SWAP1 = 0xB1 };

// Stencil neighbours of the quad c (this lane's values of the vector) from
// the other lanes, with the face values fv of load_face (or values computed
// from them).  Identical values to load_nb4 on the same vector.
template <typename T>
__device__ __forceinline__ void nb_from_lanes(const T (&c)[4], const Nb4<T>& fv, const Quad& Q,
                                              Nb4<T>& o) {
    const int lane = threadIdx.x & 63;
    const int pl = Q.li & 3;
    const int up = 4 * (lane >= 16 ? lane - 16 : lane), dn = 4 * (lane < 48 ? lane + 16 : lane);
    for (int e = 0; e < 4; ++e) {
        const T xm = lane_read(c[e], up), xp = lane_read(c[e], dn);
        const T ym = dpp_mov<DPP_ROW_SHR2>(c[e]), yp = dpp_mov<DPP_ROW_SHL2>(c[e]);
        o.xm[e] = pl == 0 ? fv.xm[e] : xm;
        o.xp[e] = pl == 3 ? fv.xp[e] : xp;
        o.ym[e] = Q.lj == 0 ? fv.ym[e] : ym;
        o.yp[e] = Q.lj == 7 ? fv.yp[e] : yp;
    }
    const T zm = dpp_mov<DPP_SWAP1>(c[3]), zp = dpp_mov<DPP_SWAP1>(c[0]);
    o.zm0 = Q.kh == 0 ? fv.zm0 : zm;
    o.zp3 = Q.kh == 1 ? fv.zp3 : zp;
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

// This workgroup's scene, the scene's workgroup range [wb, we) and the tile
// range [t0, t1) it owns (scene < 0: idle).  Consecutive logical workgroups
// sit on the same XCD so that neighbouring tiles share an L2.
struct Part {
    int b, s, wb, we, t0, t1;
};

__device__ __forceinline__ Part part_of(const PcgArgs& a) {
    Part P;
    P.b = gf::xcd_remap(blockIdx.x, a.G);
    if (a.S == 1 && !a.nact_dev) {  // one scene: the partition is b*nact/G, no table read
        P.s = 0;
        P.wb = 0;
        P.we = a.G;
        P.t0 = (int)(((long long)P.b * a.nact) / a.G);
        P.t1 = (int)(((long long)(P.b + 1) * a.nact) / a.G);
        return P;
    }
    const int4 q = *reinterpret_cast<const int4*>(a.part_info + 8 * P.b);
    const int t1 = a.part_info[8 * P.b + 4];
    P.s = __builtin_amdgcn_readfirstlane(q.x);
    P.wb = __builtin_amdgcn_readfirstlane(q.y);
    P.we = __builtin_amdgcn_readfirstlane(q.z);
    P.t0 = __builtin_amdgcn_readfirstlane(q.w);
    P.t1 = __builtin_amdgcn_readfirstlane(t1);
    return P;
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

// Workgroup sums of two values per thread with one barrier pair.
__device__ __forceinline__ void wg_sum2(double& x, double& y, double* s_red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    x = gf::wave_sum(x);
    y = gf::wave_sum(y);
    __syncthreads();
    if (lane == 0) {
        s_red[wave] = x;
        s_red[GF_WG / GF_WAVE + wave] = y;
    }
    __syncthreads();
    double tx = 0.0, ty = 0.0;
    for (int i = 0; i < GF_WG / GF_WAVE; ++i) {
        tx += s_red[i];
        ty += s_red[GF_WG / GF_WAVE + i];
    }
    x = tx;
    y = ty;
}

// This thread's share of one kind of partial over [wb, we): at most
// PCG_GMAX / GF_WG independent loads, issued together.
__device__ __forceinline__ double* pslot(double* part, int kind, int L) {
    return part + ((size_t)((L & 1) * P_SET + kind)) * PCG_GMAX;
}
__device__ __forceinline__ const double* pslot(const double* part, int kind, int L) {
    return part + ((size_t)((L & 1) * P_SET + kind)) * PCG_GMAX;
}

// This thread's share of one kind of partial written by launch L, over the
// scene's workgroups [wb, we): at most PCG_GMAX / GF_WG independent loads.
__device__ __forceinline__ double partial_share(const double* part, int kind, int L, int wb,
                                                int we) {
    const double* src = pslot(part, kind, L);
    double x[PCG_GMAX / GF_WG];
    // unconditional loads (index clamped into the slot), so all are issued
    // back to back; entries past the scene's range are dropped afterwards
#pragma unroll
    for (int q = 0; q < PCG_GMAX / GF_WG; ++q) {
        const int i = wb + threadIdx.x + q * GF_WG;
        x[q] = src[i < PCG_GMAX ? i : PCG_GMAX - 1];
    }
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < PCG_GMAX / GF_WG; ++q) {
        const int i = wb + threadIdx.x + q * GF_WG;
        v += i < we ? x[q] : 0.0;
    }
    return v;
}

__device__ __forceinline__ void put_partial2(double* part, int ka, int kb, int L, int b,
                                             double va, double vb, double* s_red) {
    wg_sum2(va, vb, s_red);
    if (threadIdx.x == 0) {
        pslot(part, ka, L)[b] = va;
        pslot(part, kb, L)[b] = vb;
    }
}

__device__ __forceinline__ void put_partial(double* part, int kind, int L, int b, double v,
                                            double* s_red) {
    const double t = wg_sum(v, s_red);
    if (threadIdx.x == 0) pslot(part, kind, L)[b] = t;
}

// Vector accessors.  The buffer number is chosen with a select: indexing the
// pointer arrays of the by-value kernel argument with a run-time value makes
// the compiler copy the argument into per-thread LDS (measured: +13 us per
// launch at 13 K tiles).
template <typename T> struct Vec;
template <> struct Vec<float> {
    static __device__ __forceinline__ float* u(const PcgArgs& a, int b) { return b ? a.u32[1] : a.u32[0]; }
    static __device__ __forceinline__ float* w(const PcgArgs& a, int b) { return b ? a.w32[1] : a.w32[0]; }
    static __device__ __forceinline__ float* s(const PcgArgs& a, int b) { return b ? a.s32[1] : a.s32[0]; }
    static __device__ __forceinline__ float* p(const PcgArgs& a) { return a.p32; }
};
template <> struct Vec<double> {
    static __device__ __forceinline__ double* u(const PcgArgs& a, int b) { return b ? a.u64[1] : a.u64[0]; }
    static __device__ __forceinline__ double* w(const PcgArgs& a, int b) { return b ? a.w64[1] : a.w64[0]; }
    static __device__ __forceinline__ double* s(const PcgArgs& a, int b) { return b ? a.s64[1] : a.s64[0]; }
    static __device__ __forceinline__ double* p(const PcgArgs& a) { return a.p64; }
};

enum { PEND_NONE = 0, PEND_SETUP = 1, PEND_REPLACE = 2 };

// A tile is handled by two whole waves, so t is wave-uniform: readfirstlane
// lets the compiler keep the tile's base and neighbour table in scalar registers.
#define FOR_QUADS(P)                                                        \
    const Quad Q = quad_of();                                               \
    const FaceOff FO = face_off(Q);                                         \
    for (int tb_ = P.t0; tb_ < P.t1; tb_ += 2)                              \
        if (const int t = __builtin_amdgcn_readfirstlane(tb_ + (threadIdx.x >> 7)); t < P.t1)

template <typename T>
__device__ __forceinline__ T zdiv(T r, uint32_t flagbyte) {
    const int d = flagbyte & 7;
    return r / (T)(d > 1 ? d : 1);
}

template <typename T>
__device__ __forceinline__ T diagm(uint32_t flagbyte) {
    const int d = flagbyte & 7;
    return (T)(d > 1 ? d : 1);
}

// p = z + beta p_old as one explicit fma, so that a cell's value is
// bit-identical whether its owner or a neighbour computes it.
template <typename T>
__device__ __forceinline__ void dir4(const T (&z)[4], const T (&po)[4], T bf, T (&pn)[4]) {
    for (int k = 0; k < 4; ++k) pn[k] = fma(bf, po[k], z[k]);
}

// Workgroup sums of three values per thread with one barrier pair.
__device__ __forceinline__ void wg_sum3(double& x, double& y, double& z, double* s_red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NW = GF_WG / GF_WAVE;
    x = gf::wave_sum(x);
    y = gf::wave_sum(y);
    z = gf::wave_sum(z);
    __syncthreads();
    if (lane == 0) {
        s_red[wave] = x;
        s_red[NW + wave] = y;
        s_red[2 * NW + wave] = z;
    }
    __syncthreads();
    double tx = 0.0, ty = 0.0, tz = 0.0;
    for (int i = 0; i < NW; ++i) {
        tx += s_red[i];
        ty += s_red[NW + i];
        tz += s_red[2 * NW + i];
    }
    x = tx;
    y = ty;
    z = tz;
}

__device__ __forceinline__ void put_partial3(double* part, int ka, int kb, int kc, int L, int b,
                                             double va, double vb, double vc, double* s_red) {
    wg_sum3(va, vb, vc, s_red);
    if (threadIdx.x == 0) {
        pslot(part, ka, L)[b] = va;
        pslot(part, kb, L)[b] = vb;
        pslot(part, kc, L)[b] = vc;
    }
}

// Logs are copied field by field: a whole-struct copy between global memory
// and a local leaves part of the local in per-thread LDS.
#define GF_LOG_FIELDS(X)                                                              \
    X(rz) X(rz_old) X(rr) X(bb) X(alpha) X(alpha2) X(rr_upd) X(rr_true) X(pap) X(iters) \
    X(done) X(converged) X(maxed) X(upd) X(merge) X(nupd) X(pending) X(fresh) X(pbuf) X(xpend) \
    X(alpha3)

__device__ __forceinline__ PcgLog copy_log(const PcgLog* src) {
    PcgLog e;
#define GF_LOG_LD(f) e.f = src->f;
    GF_LOG_FIELDS(GF_LOG_LD)
#undef GF_LOG_LD
    e.pad_i = 0;
    return e;
}
__device__ __forceinline__ void put_log(PcgLog* dst, const PcgLog& e) {
#define GF_LOG_ST(f) dst->f = e.f;
    GF_LOG_FIELDS(GF_LOG_ST)
#undef GF_LOG_ST
}

__device__ __forceinline__ PcgLog live_log(const PcgArgs& a, int s, int L) {
    return copy_log(a.log + 2 * s + (L & 1));
}
__device__ __forceinline__ void set_live_log(const PcgArgs& a, int s, int L, const PcgLog& e) {
    put_log(a.log + 2 * s + (L & 1), e);
}

__device__ __forceinline__ void write_log(const PcgArgs& a, const Part& P, const PcgLog& e,
                                          int L) {
    if (P.b == P.wb && threadIdx.x == 0) {
        set_live_log(a, P.s, L, e);
        if (a.hout) put_log(a.hout + 2 * P.s + (L & 1), e);
    }
}

// ---- the replacement's last workgroup --------------------------------------
// The two-kernel form's replacement finishes in the scene's last workgroup
// to arrive instead of a separate one-workgroup launch.  Its partials are
// written and read with device-coherent accesses (cache policy sc1, past the
// caches that are not coherent between XCDs), and each workgroup's stores are
// acknowledged before it counts itself, so no cache write-back or
// invalidation (device-scope fences cost about 20 us over 1,024 workgroups on
// MI355X) is needed.  No workgroup waits for another.  Counters: one per
// scene in a spare partial kind (set 0), zeroed by the partition kernel and
// reset by the last workgroup.
constexpr int P_ARRIVE = P_KINDS;
static_assert(P_ARRIVE < P_SET, "arrival counters need a spare partial kind");
__device__ __forceinline__ unsigned* arrive_cnt(double* part) {
    return reinterpret_cast<unsigned*>(pslot(part, P_ARRIVE, 0));
}
__device__ __forceinline__ void st_coh(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_coh(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// partial_share with device-coherent loads.
__device__ __forceinline__ double partial_share_coh(const double* part, int kind, int L, int wb,
                                                    int we) {
    const double* src = pslot(part, kind, L);
    double x[PCG_GMAX / GF_WG];
#pragma unroll
    for (int q = 0; q < PCG_GMAX / GF_WG; ++q) {
        const int i = wb + threadIdx.x + q * GF_WG;
        x[q] = ld_coh(src + (i < PCG_GMAX ? i : PCG_GMAX - 1));
    }
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < PCG_GMAX / GF_WG; ++q) {
        const int i = wb + threadIdx.x + q * GF_WG;
        v += i < we ? x[q] : 0.0;
    }
    return v;
}

// put_partial2 with device-coherent stores, then the arrival count: true (in
// every thread) in the scene's last workgroup to arrive.
__device__ __forceinline__ bool put_partial2_last(double* part, int ka, int kb, int L,
                                                  const Part& P, double va, double vb,
                                                  double* s_red, int* s_last) {
    wg_sum2(va, vb, s_red);
    if (threadIdx.x == 0) {
        st_coh(pslot(part, ka, L) + P.b, va);
        st_coh(pslot(part, kb, L) + P.b, vb);
        // the partials are acknowledged before the workgroup counts itself
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        unsigned* c = arrive_cnt(part) + P.s;
        const bool last = atomicAdd(c, 1u) + 1u == (unsigned)(P.we - P.wb);
        if (last) atomicExch(c, 0u);
        *s_last = last;
    }
    __syncthreads();
    return *s_last != 0;
}

// The end of a residual replacement from the scene's sums: the true residual
// r.r, the new gamma (r.z) and delta (one-kernel form; zero in the two-kernel
// form); the solve stops when the true residual meets the tolerance or the
// iterations are spent.
__device__ __forceinline__ void finish_replace(const PcgArgs& a, PcgLog& e, double rrt,
                                               double gam, double del) {
    e.rr_true = rrt; e.rr = rrt; e.rr_upd = rrt; e.rz = gam; e.pap = del;
    e.upd = 0; e.merge = 1; e.nupd += 1;
    if (rrt < a.tol2 * e.bb) { e.converged = 1; e.done = 1; }
    else if (e.iters >= a.max_iter) { e.maxed = 1; e.done = 1; }
    else e.done = 0;
    e.pending = PEND_NONE;
}

// Consumes the last iteration's partials (when fresh): gamma = (r,u),
// delta = (Au,u), r.r; counts the finished iteration and sets the stop /
// replacement flags.  Identical in every workgroup of the scene.
template <bool MIXED>
__device__ __forceinline__ void consume(const PcgArgs& a, PcgLog& e, double gam, double del,
                                        double rr, double* s_red) {
    if (!e.fresh) return;
    wg_sum3(gam, del, rr, s_red);
    e.rz_old = e.rz;
    e.rz = gam;
    e.pap = del;
    e.rr = rr;
    e.fresh = 0;
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
}

// The previous launch's partial sums do not depend on the solver log, but
// they are only used past the log's branches, where the compiler would sink
// their loads and pay two memory latencies in a row.  This empty asm uses
// them together with the log, so both sets of loads are in flight at once.
__device__ __forceinline__ void issue_together(const PcgLog& e, double x, double y, double z) {
    asm volatile("" ::"v"(x), "v"(y), "v"(z), "s"(e.done), "s"(e.fresh));
}

__device__ __forceinline__ bool pass_through(const PcgArgs& a, const Part& P, const PcgLog& e,
                                             int L) {
    if (!e.done) return false;
    write_log(a, P, e, L);
    return true;
}

// Workgroup partition of a solve (one workgroup): scene s gets
// 1 + floor((G - NE) * tiles_s / tiles) workgroups when it has active tiles
// (NE = scenes with tiles), none otherwise; each workgroup of a scene takes an
// even share of its tiles.  A single scene gets at most one workgroup per
// two tiles, which is the grid the host-count form launches, so both forms
// split the tiles (and order the partial sums) identically.  Also marks every
// scene's log as awaiting the setup.
__global__ __launch_bounds__(GF_WG) void k_pcg_partition(PcgArgs a) {
    __shared__ int s_wave[GF_WG / GF_WAVE + 1];
    __shared__ double s_red[2 * GF_WG / GF_WAVE];
    int32_t* wg_off = a.part_info + 8 * PCG_GMAX;
    const int n_one = a.nact_dev ? a.nact_dev[0] : a.nact;  // tiles of a single scene
    long long tiles = 0, ne = 0;
    for (int sc = threadIdx.x; sc < a.S; sc += GF_WG) {
        const int n = a.scene_slot ? a.scene_slot[sc + 1] - a.scene_slot[sc] : n_one;
        tiles += n;
        ne += n > 0;
    }
    tiles = (long long)wg_sum((double)tiles, s_red);  // integer sums, exact in FP64
    ne = (long long)wg_sum((double)ne, s_red);
    const long long spare = a.G - ne > 0 ? a.G - ne : 0;
    int carry = 0;
    for (int base = 0; base < a.S; base += GF_WG) {
        const int sc = base + threadIdx.x;
        int g = 0, n = 0, slot0 = 0;
        if (sc < a.S) {
            slot0 = a.scene_slot ? a.scene_slot[sc] : 0;
            n = a.scene_slot ? a.scene_slot[sc + 1] - slot0 : n_one;
            g = n > 0 ? 1 + (int)((spare * n) / (tiles > 0 ? tiles : 1)) : 0;
            if (!a.scene_slot && g > (n + 1) / 2) g = (n + 1) / 2;
        }
        int total;
        const int off = carry + gf::wg_exclusive_scan(g, s_wave, &total);
        if (sc < a.S) {
            wg_off[sc] = off;
            PcgLog e = {};
            e.pending = PEND_SETUP;
            put_log(a.log + 2 * sc, e);
            arrive_cnt(a.part)[sc] = 0u;
        }
        carry += total;
    }
    if (threadIdx.x == 0) wg_off[a.S] = carry;
    __syncthreads();
    // one entry per workgroup, filled in parallel: its scene is the last one
    // whose first workgroup is at or before it
    for (int b = threadIdx.x; b < a.G; b += GF_WG) {
        int32_t* w = a.part_info + 8 * b;
        if (b >= carry) {
            w[0] = -1; w[1] = 0; w[2] = 0; w[3] = 0; w[4] = 0;
            continue;
        }
        int lo = 0, hi = a.S - 1;  // wg_off[lo] <= b < wg_off[hi + 1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wg_off[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        const int sc = lo, off = wg_off[sc], g = wg_off[sc + 1] - off, q = b - off;
        const int slot0 = a.scene_slot ? a.scene_slot[sc] : 0;
        const int n = a.scene_slot ? a.scene_slot[sc + 1] - slot0 : n_one;
        w[0] = sc; w[1] = off; w[2] = off + g;
        w[3] = slot0 + (int)(((long long)q * n) / g);
        w[4] = slot0 + (int)(((long long)(q + 1) * n) / g);
    }
}

// y = A v on a quad (fluid cells; 0 elsewhere) given its neighbour values.
template <typename T>
__device__ __forceinline__ void apply_A4(const T (&c)[4], const Nb4<T>& o, uint32_t fw,
                                         T (&y)[4]) {
    for (int k = 0; k < 4; ++k)
        y[k] = keep_if((T)fl_diag(fw, k) * c[k] - nb_sum4(o, c, k), fl_fluid(fw, k));
}

// Setup, pass 1: x = 0, p = 0, s~ = 0, u = M^-1 (rhs_sign * b) on fluid cells.
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_cg1_setup(PcgArgs a) {
    const Part P = part_of(a);
    if (P.s < 0) return;
    T* u = Vec<T>::u(a, 0);
    T* sv = Vec<T>::s(a, 0);
    T* p = Vec<T>::p(a);
    FOR_QUADS(P) {
        const size_t idx = (size_t)t * GF_TILE_CELLS + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        float bv[4];
        ld4(a.b + idx, bv);
        T uv[4], zero[4] = {(T)0, (T)0, (T)0, (T)0};
        double zd[4] = {0.0, 0.0, 0.0, 0.0};
        for (int e = 0; e < 4; ++e) {
            const T rv = fl_fluid(fw, e) ? (T)(a.rhs_sign * bv[e]) : (T)0;
            uv[e] = zdiv(rv, fw >> (8 * e));
        }
        st4(u + idx, uv);
        st4(p + idx, zero);
        st4(sv + idx, zero);
        st4(a.x64 + idx, zd);
        if (MIXED) {
            float zf[4] = {0.f, 0.f, 0.f, 0.f};
            st4(a.x32 + idx, zf);
        }
    }
}

// Setup, pass 2: w~ = M^-1 A u; partials gamma0 = (r,u), delta0 = (Au,u), b.b.
template <typename T>
__global__ __launch_bounds__(GF_WG) void k_cg1_setup2(PcgArgs a) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    const T* __restrict__ u = Vec<T>::u(a, 0);
    T* __restrict__ w = Vec<T>::w(a, 0);
    double gam = 0.0, del = 0.0, rr = 0.0;
    FOR_QUADS(P) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        T c[4], y[4], wv[4];
        ld4(u + idx, c);
        Nb4<T> fu, o;
        load_face(u, face_tiles(a.nbr, t), FO, fu);
        nb_from_lanes(c, fu, Q, o);
        apply_A4(c, o, fw, y);
        for (int k = 0; k < 4; ++k) {
            const T dm = diagm<T>(fw >> (8 * k));
            wv[k] = y[k] / dm;
            const double r = (double)dm * (double)c[k];
            gam += r * (double)c[k];
            del += (double)y[k] * (double)c[k];
            rr += r * r;
        }
        st4(w + idx, wv);
    }
    put_partial3(a.part, P_RZ, P_PAP, P_BB, 0, P.b, gam, del, rr, s_red);  // read by fin at L = 1
}

// Finalisation after the setup: one workgroup per scene (grid S).  (A
// residual replacement finishes in its own last workgroup.)
__global__ __launch_bounds__(GF_WG) void k_pcg_fin(PcgArgs a, int L) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const int sc = blockIdx.x;
    const int32_t* wg_off = a.part_info + 8 * PCG_GMAX;
    const int wb = wg_off[sc], we = wg_off[sc + 1];
    PcgLog e = live_log(a, sc, L - 1);
    if (e.pending == PEND_SETUP) {  // setup2 wrote launch 0's set; fin runs at L = 1
        double gam = partial_share(a.part, P_RZ, L - 1, wb, we);
        double del = partial_share(a.part, P_PAP, L - 1, wb, we);
        double bb = partial_share(a.part, P_BB, L - 1, wb, we);
        wg_sum3(gam, del, bb, s_red);
        e.rz = gam; e.pap = del; e.rz_old = 0.0; e.alpha = 0.0;
        e.bb = bb; e.rr = bb; e.rr_upd = bb; e.rr_true = bb;
        e.iters = 0; e.maxed = 0; e.upd = 0; e.merge = 0; e.nupd = 0; e.fresh = 0; e.pbuf = 0;
        e.xpend = 0; e.alpha2 = 0.0; e.alpha3 = 0.0;
        const int zero = bb < 1e-30;
        e.converged = zero;
        e.done = zero || a.max_iter <= 0;
    }
    e.pending = PEND_NONE;
    if (threadIdx.x == 0) {
        set_live_log(a, sc, L, e);
        if (a.hout) put_log(a.hout + 2 * sc + (L & 1), e);
    }
}

// u_new = u - alpha (w~ + beta s~) for a quad's neighbour values (the same
// two fmas as the owner uses, so both sides agree bit for bit).
template <typename T>
__device__ __forceinline__ void unew4(const T (&u)[4], const T (&w)[4], const T (&sv)[4], T al,
                                      T be, T (&o)[4]) {
    for (int k = 0; k < 4; ++k) o[k] = fma(-al, fma(be, sv[k], w[k]), u[k]);
}

// One iteration (see the header): reads u, w~, s~ of buffer e.pbuf (own
// cells and stencil neighbours), p and x; writes p, x in place and u, w~, s~
// to the other buffer; partials of the next gamma, delta and r.r.
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_cg1_iter(PcgArgs a, int L) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    PcgLog e = live_log(a, P.s, L - 1);
    const double gsh = partial_share(a.part, P_RZ, L - 1, P.wb, P.we);
    const double dsh = partial_share(a.part, P_PAP, L - 1, P.wb, P.we);
    const double rrs = partial_share(a.part, P_RR, L - 1, P.wb, P.we);
    issue_together(e, gsh, dsh, rrs);
    if (pass_through(a, P, e, L)) return;
    consume<MIXED>(a, e, gsh, dsh, rrs, s_red);
    if (e.done) {
        write_log(a, P, e, L);
        return;
    }
    const double beta = (e.iters == 0 || fabs(e.rz_old) < 1e-30) ? 0.0 : e.rz / e.rz_old;
    double den = e.pap;
    if (beta != 0.0 && e.alpha != 0.0) den -= beta * e.rz / e.alpha;
    const double alpha = fabs(den) < 1e-30 ? 0.0 : e.rz / den;
    const T be = (T)beta, al = (T)alpha;
    const int cur = e.pbuf, nxt = e.pbuf ^ 1;
    const T* __restrict__ u = Vec<T>::u(a, cur);
    const T* __restrict__ w = Vec<T>::w(a, cur);
    const T* __restrict__ sv = Vec<T>::s(a, cur);
    T* __restrict__ un = Vec<T>::u(a, nxt);
    T* __restrict__ wn = Vec<T>::w(a, nxt);
    T* __restrict__ sn = Vec<T>::s(a, nxt);
    T* __restrict__ p = Vec<T>::p(a);
    const int merge = MIXED ? e.merge : 0;
    double gam = 0.0, del = 0.0, rr = 0.0;
    FOR_QUADS(P) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const FaceTiles ft = face_tiles(a.nbr, t);
        const uint32_t fw = ld_flags(a.flags + base, FO.own);
        T uq[4], wq[4], sq[4], pq[4];
        ld4(u + base + FO.own, uq);
        ld4(w + base + FO.own, wq);
        ld4(sv + base + FO.own, sq);
        ld4(p + base + FO.own, pq);
        Nb4<T> nu, nw, ns;
        load_face(u, ft, FO, nu);
        load_face(w, ft, FO, nw);
        load_face(sv, ft, FO, ns);
        // own cells
        T pn[4], snew[4], c[4];
        for (int k = 0; k < 4; ++k) {
            pn[k] = fma(be, pq[k], uq[k]);
            snew[k] = fma(be, sq[k], wq[k]);
            c[k] = fma(-al, snew[k], uq[k]);
        }
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
            for (int k = 0; k < 4; ++k) xo[k] = xo[k] + (float)al * (float)pn[k];
            st4(a.x32 + idx, xo);
        } else {
            double xd[4];
            ld4(a.x64 + idx, xd);
            for (int k = 0; k < 4; ++k) xd[k] += (double)al * (double)pn[k];
            st4(a.x64 + idx, xd);
        }
        // face cells' new u (the owner's two fmas, bit for bit); the rest of
        // the stencil from the lanes' own new u
        Nb4<T> fc, o;
        unew4(nu.xm, nw.xm, ns.xm, al, be, fc.xm);
        unew4(nu.xp, nw.xp, ns.xp, al, be, fc.xp);
        unew4(nu.ym, nw.ym, ns.ym, al, be, fc.ym);
        unew4(nu.yp, nw.yp, ns.yp, al, be, fc.yp);
        fc.zm0 = fma(-al, fma(be, ns.zm0, nw.zm0), nu.zm0);
        fc.zp3 = fma(-al, fma(be, ns.zp3, nw.zp3), nu.zp3);
        nb_from_lanes(c, fc, Q, o);
        T y[4], wv[4];
        apply_A4(c, o, fw, y);
        for (int k = 0; k < 4; ++k) {
            const T dm = diagm<T>(fw >> (8 * k));
            wv[k] = y[k] / dm;
            const double r = (double)dm * (double)c[k];
            gam += r * (double)c[k];
            del += (double)y[k] * (double)c[k];
            rr += r * r;
        }
        st4(p + idx, pn);
        st4(un + idx, c);
        st4(wn + idx, wv);
        st4(sn + idx, snew);
    }
    put_partial3(a.part, P_RZ, P_PAP, P_RR, L, P.b, gam, del, rr, s_red);
    e.alpha = alpha;
    e.pbuf = nxt;
    e.merge = 0;
    e.fresh = 1;
    write_log(a, P, e, L);
}

// Chunk end: consumes the last iteration's partials and decides.  In mixed
// mode, when a replacement is due or convergence is to be confirmed, pass 1
// of the replacement: the FP64 true residual r = rhs - A (x64 + x32), its
// FP32 u = M^-1 r, and s~ = M^-1 A p, into the current buffers; partial of the
// true r.r.
template <bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_cg1_replace(PcgArgs a, int L) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    PcgLog e = live_log(a, P.s, L - 1);
    const double gsh = partial_share(a.part, P_RZ, L - 1, P.wb, P.we);
    const double dsh = partial_share(a.part, P_PAP, L - 1, P.wb, P.we);
    const double rrs = partial_share(a.part, P_RR, L - 1, P.wb, P.we);
    issue_together(e, gsh, dsh, rrs);
    consume<MIXED>(a, e, gsh, dsh, rrs, s_red);
    const bool run = MIXED && (e.upd || (e.done && !e.converged));
    if (!run) {
        write_log(a, P, e, L);
        return;
    }
    float* __restrict__ u = a.u32[e.pbuf];
    float* __restrict__ sv = a.s32[e.pbuf];
    double rrt = 0.0;
    FOR_QUADS(P) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        const FaceTiles ft = face_tiles(a.nbr, t);
        double cd[4];
        float cf[4], bv[4], pq[4];
        ld4(a.x64 + idx, cd);
        ld4(a.x32 + idx, cf);
        ld4(a.b + idx, bv);
        ld4(a.p32 + idx, pq);
        Nb4<double> fd, od;
        Nb4<float> ff, fp, op;
        load_face(a.x64, ft, FO, fd);
        load_face(a.x32, ft, FO, ff);
        load_face(a.p32, ft, FO, fp);
        double c[4];
        for (int k = 0; k < 4; ++k) {
            c[k] = cd[k] + (double)cf[k];
            fd.xm[k] += (double)ff.xm[k];
            fd.xp[k] += (double)ff.xp[k];
            fd.ym[k] += (double)ff.ym[k];
            fd.yp[k] += (double)ff.yp[k];
        }
        fd.zm0 += (double)ff.zm0;
        fd.zp3 += (double)ff.zp3;
        nb_from_lanes(c, fd, Q, od);
        nb_from_lanes(pq, fp, Q, op);
        float ap[4], uo[4], so[4];
        apply_A4(pq, op, fw, ap);
        for (int k = 0; k < 4; ++k) {
            const int d = fl_diag(fw, k), fl = fl_fluid(fw, k);
            const double ax = (double)d * c[k] - nb_sum4(od, c, k);
            const double rv = keep_if((double)(a.rhs_sign * bv[k]) - ax, fl);
            const float dm = (float)(d > 1 ? d : 1);
            uo[k] = (float)rv / dm;
            so[k] = ap[k] / dm;
            rrt += rv * rv;
        }
        st4(u + idx, uo);
        st4(sv + idx, so);
    }
    put_partial(a.part, P_RRT, L, P.b, rrt, s_red);
    e.pending = PEND_REPLACE;
    write_log(a, P, e, L);
}

// Pass 2 of a replacement (mixed mode; passes the log through otherwise):
// w~ = M^-1 A u from the new u; partials of gamma and delta, which the
// scene's last workgroup to arrive sums with pass 1's r.r to finish the log.
__global__ __launch_bounds__(GF_WG) void k_cg1_replace2(PcgArgs a, int L) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    PcgLog e = live_log(a, P.s, L - 1);
    if (e.pending != PEND_REPLACE) {
        write_log(a, P, e, L);
        return;
    }
    const float* __restrict__ u = a.u32[e.pbuf];
    float* __restrict__ w = a.w32[e.pbuf];
    double gam = 0.0, del = 0.0;
    FOR_QUADS(P) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        float c[4], y[4], wv[4];
        ld4(u + idx, c);
        Nb4<float> fu, o;
        load_face(u, face_tiles(a.nbr, t), FO, fu);
        nb_from_lanes(c, fu, Q, o);
        apply_A4(c, o, fw, y);
        for (int k = 0; k < 4; ++k) {
            const float dm = diagm<float>(fw >> (8 * k));
            wv[k] = y[k] / dm;
            gam += (double)dm * (double)c[k] * (double)c[k];
            del += (double)y[k] * (double)c[k];
        }
        st4(w + idx, wv);
    }
    __shared__ int s_last;
    if (!put_partial2_last(a.part, P_RZT, P_BB, L, P, gam, del, s_red, &s_last)) return;
    // the scene's last workgroup: pass 1's true residual, gamma and delta
    double rrt = partial_share_coh(a.part, P_RRT, L - 1, P.wb, P.we);
    gam = partial_share_coh(a.part, P_RZT, L, P.wb, P.we);
    del = partial_share_coh(a.part, P_BB, L, P.wb, P.we);
    wg_sum3(rrt, gam, del, s_red);
    finish_replace(a, e, rrt, gam, del);
    if (threadIdx.x == 0) {
        set_live_log(a, P.s, L, e);
        if (a.hout) put_log(a.hout + 2 * P.s + (L & 1), e);
    }
}

// ---- two-kernel (standard) PCG -----------------------------------------
//   dirspmv: beta = rz/rz_old, p = z + beta p (into the next slot of a ring
//            of four direction buffers; the stencil neighbours' p is
//            recomputed from their z and old p), p.Ap
//   update:  alpha = rz/pAp, Ap again from the stored p (bit-identical to the
//            direction kernel's), x += alpha p (every few updates, for the
//            pending terms at once), r = M z - alpha Ap, z = r/M, r.z, r.r
// Ap is recomputed rather than stored: the stencil re-reads p through the
// caches, while storing and reloading Ap would cost 8 bytes per cell of HBM
// traffic per iteration.  Deferring x's terms saves the x read and write
// (8 or 16 bytes per cell) of the deferring updates; the flushing update
// reads the pending directions instead.  Storage: z in u[0], the direction
// ring in {p, s[0], u[1], w[1]}, the packed flags in w[0].

template <typename T>
__device__ __forceinline__ T* std_p(const PcgArgs& a, int buf) {
    // a ring of four direction buffers: p, s~[0] and the one-kernel form's
    // second u and w~ (this form does not use them otherwise)
    buf &= 3;
    return buf == 0 ? Vec<T>::p(a) : (buf == 1 ? Vec<T>::s(a, 0) : (buf == 2 ? Vec<T>::u(a, 1) : Vec<T>::w(a, 1)));
}

// The two-kernel form adds alpha p to the iterate x on every XE-th update
// (and on a chunk's last one); the updates in between leave their terms
// pending.  xmode argument of the update kernel: X_DEFER, or the number of
// pending terms the host expects this update to add besides its own.
constexpr int X_DEFER = -1;
constexpr int X_MAX_PEND = 3;  // pending terms; the ring holds them and the current p

// The alpha of pending term i (0 = the latest) in a log.
__device__ __forceinline__ double pend_alpha(const PcgLog& e, int i) {
    return i == 0 ? e.alpha : (i == 1 ? e.alpha2 : e.alpha3);
}

// x += a[n-1] p[n-1] + ... + a[0] p[0] on a quad, term by term in iteration
// order (oldest first), as one update per iteration would add them, so
// deferring a term changes no bit of x (mixed: x32, first folded into x64 when
// merge is set).  The direction kernel's flush of a solve that stopped with
// terms pending; the update kernel adds in the same order (upd_compute).
template <typename T, bool MIXED>
__device__ __forceinline__ void x_addn(const PcgArgs& a, size_t idx, int merge, int n,
                                       const T (&al)[X_MAX_PEND],
                                       const T (&pp)[X_MAX_PEND][4]) {
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
#pragma unroll
        for (int i = X_MAX_PEND - 1; i >= 0; --i)
            if (i < n)
                for (int k = 0; k < 4; ++k) xo[k] = xo[k] + (float)al[i] * (float)pp[i][k];
        st4(a.x32 + idx, xo);
    } else {
        double xd[4];
        ld4(a.x64 + idx, xd);
#pragma unroll
        for (int i = X_MAX_PEND - 1; i >= 0; --i)
            if (i < n)
                for (int k = 0; k < 4; ++k) xd[k] += (double)al[i] * (double)pp[i][k];
        st4(a.x64 + idx, xd);
    }
}

// The update kernel's r.z / r.r partials (written by launch L - 1).
template <bool MIXED>
__device__ __forceinline__ void consume_std(const PcgArgs& a, PcgLog& e, double rz_new, double rr,
                                            double* s_red) {
    if (!e.fresh) return;
    wg_sum2(rz_new, rr, s_red);
    e.rz_old = e.rz;
    e.rz = rz_new;
    e.rr = rr;
    e.fresh = 0;
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
}

// Setup: x = 0, z = M^-1 (rhs_sign*b), p (buffer 0) = 0; partials r.z and b.b
// (and a zero delta, so k_pcg_fin reads both forms alike).
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_setup(PcgArgs a) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    T* z = Vec<T>::u(a, 0);
    T* p = std_p<T>(a, 0);
    double rz = 0.0, rr = 0.0;
    FOR_QUADS(P) {
        const size_t idx = (size_t)t * GF_TILE_CELLS + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        a.pflags[idx >> 2] = (uint16_t)pack_flags(fw);
        float bv[4];
        ld4(a.b + idx, bv);
        T zv[4], zero[4] = {(T)0, (T)0, (T)0, (T)0};
        double zd[4] = {0.0, 0.0, 0.0, 0.0};
        for (int e = 0; e < 4; ++e) {
            const T rv = fl_fluid(fw, e) ? (T)(a.rhs_sign * bv[e]) : (T)0;
            zv[e] = zdiv(rv, fw >> (8 * e));
            rz += (double)rv * (double)zv[e];
            rr += (double)rv * (double)rv;
        }
        st4(z + idx, zv);
        st4(p + idx, zero);
        st4(a.x64 + idx, zd);
        if (MIXED) {
            float zf[4] = {0.f, 0.f, 0.f, 0.f};
            st4(a.x32 + idx, zf);
        }
    }
    put_partial3(a.part, P_RZ, P_PAP, P_BB, 0, P.b, rz, 0.0, rr, s_red);  // read by fin at L = 1
}

// Loads of one tile pass of the direction kernel: this quad's z and old p,
// and both across the +x, +y and +z faces of the wave's block.
template <typename T>
struct DirIn {
    uint32_t fw;
    T zq[4], pq[4], zx[4], px[4], zy[4], py[4];
    T zz, pz;
};

// t < 0: no tile (reads the zero tile; the values are not used).  row is the
// tile's stored neighbour row (face_row), read one pass ahead.
template <typename T>
__device__ __forceinline__ void dir_issue(DirIn<T>& d, const PcgArgs& a, const T* __restrict__ z,
                                          const T* __restrict__ po, int t, const FaceTiles& row,
                                          const FaceOff& FO) {
    const FaceTiles ft = for_half(row, t);
    d.fw = ld_pflags(a.pflags, t, FO.own);
    ld4(tile_base(z, t) + FO.own, d.zq);
    ld4(tile_base(po, t) + FO.own, d.pq);
    ld4(tile_base(z, ft.xp) + FO.xp, d.zx);
    ld4(tile_base(po, ft.xp) + FO.xp, d.px);
    ld4(tile_base(z, ft.yp) + FO.yp, d.zy);
    ld4(tile_base(po, ft.yp) + FO.yp, d.py);
    d.zz = tile_base(z, ft.zp)[FO.zp];
    d.pz = tile_base(po, ft.zp)[FO.zp];
}

// p = z + beta p_old for the quad (stored into pn) and p.Ap as the quadratic
// form sum_i d_i p_i^2 - 2 sum_{edges} p_i p_j (A symmetric, p zero off fluid
// cells): each cell with its +x, +y, +z neighbours, products exact in FP64.
// The face cells' new p comes from their z and old p with the owner's fma,
// bit for bit.
template <typename T>
__device__ __forceinline__ void dir_compute(const DirIn<T>& d, T bf, const Quad& Q, int dn,
                                            T* __restrict__ pn, size_t idx, double& pap) {
    T c[4], cx[4], cy[4];
    dir4(d.zq, d.pq, bf, c);
    dir4(d.zx, d.px, bf, cx);
    dir4(d.zy, d.py, bf, cy);
    const T cz = fma(bf, d.pz, d.zz);
    const T zp = dpp_mov<DPP_SWAP1>(c[0]);
    for (int k = 0; k < 4; ++k) {
        // every lane takes part in the exchange (the sources are other lanes)
        const T lx = lane_read(c[k], dn), ly = dpp_mov<DPP_ROW_SHL2>(c[k]);
        const T nxp = (Q.li & 3) == 3 ? cx[k] : lx;
        const T nyp = Q.lj == 7 ? cy[k] : ly;
        const T nzp = k < 3 ? c[k + 1] : (Q.kh == 1 ? cz : zp);
        const double cd = (double)c[k];
        const double nb = (double)nxp + (double)nyp + (double)nzp;
        pap += cd * ((double)fl_diag(d.fw, k) * cd - 2.0 * nb);
    }
    st4(pn + idx, c);
}

// Every load of a tile pass is issued before any of its values is used (one
// memory round trip per pass), and the first pass's loads go out before the
// solver log and the partial sums are waited for.  They read the direction
// buffer pb that the host expects from the launch sequence; a scene whose log
// names the other buffer reloads them.
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_dirspmv(PcgArgs a, int L, int pb) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    const Quad Q = quad_of();
    const FaceOff FO = face_off(Q);
    const int half = threadIdx.x >> 7;
    const T* __restrict__ z = Vec<T>::u(a, 0);
    const bool any = P.t0 < P.t1;
    const int tf = __builtin_amdgcn_readfirstlane(any ? min(P.t0 + half, P.t1 - 1) : -1);
    // unconditional, so that the wait counts below see one path
    const FaceTiles ftf = face_row(a.nbr, tf);
    DirIn<T> d;
    dir_issue(d, a, z, std_p<T>(a, pb), tf, ftf, FO);
    const double rzs = partial_share(a.part, P_RZ, L - 1, P.wb, P.we);
    const double rrs = partial_share(a.part, P_RR, L - 1, P.wb, P.we);
    PcgLog e = live_log(a, P.s, L - 1);
    issue_together(e, rzs, rrs, 0.0);
    if (pass_through(a, P, e, L)) return;
    consume_std<MIXED>(a, e, rzs, rrs, s_red);
    if (e.done) {
        if (e.xpend) {  // the solve stopped with alpha p of the last updates pending
            const int n = e.xpend;
            T al[X_MAX_PEND], pp[X_MAX_PEND][4];
#pragma unroll
            for (int i = 0; i < X_MAX_PEND; ++i) al[i] = (T)pend_alpha(e, i);
            FOR_QUADS(P) {
                const size_t idx = (size_t)t * GF_TILE_CELLS + FO.own;
#pragma unroll
                for (int i = 0; i < X_MAX_PEND; ++i)
                    if (i < n) ld4(std_p<T>(a, e.pbuf - i) + idx, pp[i]);
                x_addn<T, MIXED>(a, idx, MIXED ? e.merge : 0, n, al, pp);
            }
            e.xpend = 0;
            e.merge = 0;
        }
        write_log(a, P, e, L);
        return;
    }
    const double beta = fabs(e.rz_old) < 1e-30 ? 0.0 : e.rz / e.rz_old;
    const T bf = (T)beta;
    const T* __restrict__ po = std_p<T>(a, e.pbuf);
    T* __restrict__ pn = std_p<T>(a, e.pbuf + 1);
    if (any && e.pbuf != pb) dir_issue(d, a, z, po, tf, ftf, FO);
    double pap = 0.0;
    const int lane = threadIdx.x & 63;
    const int dn = 4 * (lane < 48 ? lane + 16 : lane);  // +x: sixteen lanes up
    if (any) {
        int tn = __builtin_amdgcn_readfirstlane(P.t0 + 2 + half);
        FaceTiles ftn = face_row(a.nbr, tn < P.t1 ? tn : -1);
        if (P.t0 + half < P.t1) dir_compute(d, bf, Q, dn, pn, (size_t)tf * GF_TILE_CELLS + FO.own, pap);
        for (int tb = P.t0 + 2; tb < P.t1; tb += 2) {
            const int t = tn;
            if (t >= P.t1) break;
            DirIn<T> n;
            dir_issue(n, a, z, po, t, ftn, FO);
            // the next pass's row is read after this pass's vector loads are out
            __builtin_amdgcn_sched_barrier(0);
            tn = __builtin_amdgcn_readfirstlane(tb + 2 + half);
            ftn = face_row(a.nbr, tn < P.t1 ? tn : -1);
            dir_compute(n, bf, Q, dn, pn, (size_t)t * GF_TILE_CELLS + FO.own, pap);
        }
    }
    put_partial(a.part, P_PAP, L, P.b, pap, s_red);
    e.pbuf = (e.pbuf + 1) & 3;
    write_log(a, P, e, L);
}

// Loads of one tile pass of the update kernel: flags, p, z, the iterate x
// and the pending terms' directions (when this update adds to x) and p across
// the six faces of the wave's block.
template <typename T, bool MIXED>
struct UpdIn {
    using XT = typename std::conditional<MIXED, float, double>::type;
    uint32_t fw;
    T pv[4], zv[4], pp[X_MAX_PEND][4];
    XT xo[4];
    Nb4<T> fp;
};

// np >= 0: this update adds to x, with np pending terms whose directions sit
// in the ring buffers before pb; np < 0: it defers.
template <typename T, bool MIXED>
__device__ __forceinline__ void upd_issue(UpdIn<T, MIXED>& d, const PcgArgs& a, int pb,
                                          const T* __restrict__ zp, int t, const FaceTiles& row,
                                          const FaceOff& FO, int np) {
    // t < 0: no tile (reads the zero tile; the values are not used); row: the
    // stored neighbour row (face_row), read one pass ahead
    const FaceTiles ft = for_half(row, t);
    const uint32_t o = FO.own;
    const T* __restrict__ p = std_p<T>(a, pb);
    d.fw = ld_pflags(a.pflags, t, o);
    ld4(tile_base(p, t) + o, d.pv);
    ld4(tile_base(zp, t) + o, d.zv);
    load_face(p, ft, FO, d.fp);
    if (np >= 0) {
        if (MIXED) ld4(tile_base(a.x32, t) + o, reinterpret_cast<float(&)[4]>(d.xo));
        else ld4(tile_base(a.x64, t) + o, reinterpret_cast<double(&)[4]>(d.xo));
    }
#pragma unroll
    for (int i = 0; i < X_MAX_PEND; ++i)
        if (i < np) ld4(tile_base(std_p<T>(a, pb - 1 - i), t) + o, d.pp[i]);
}

// Ap from p (the direction kernel's operands and order), x += (pending
// terms) + alpha p, r = M z - alpha Ap, z = r / M, and the r.z, r.r partials.
template <typename T, bool MIXED>
__device__ __forceinline__ void upd_compute(UpdIn<T, MIXED>& d, const PcgArgs& a, const Quad& Q,
                                            T af, const T (&al)[X_MAX_PEND], int np, int merge,
                                            T* __restrict__ zp, size_t idx, double& rz,
                                            double& rr) {
    T apv[4];
    Nb4<T> nq;
    nb_from_lanes(d.pv, d.fp, Q, nq);
    apply_A4(d.pv, nq, d.fw, apv);
    if (np >= 0) {
        if (MIXED) {
            float(&xo)[4] = reinterpret_cast<float(&)[4]>(d.xo);
            if (merge) {
                double xd[4];
                ld4(a.x64 + idx, xd);
                for (int k = 0; k < 4; ++k) {
                    xd[k] += (double)xo[k];
                    xo[k] = 0.0f;
                }
                st4(a.x64 + idx, xd);
            }
#pragma unroll
            for (int i = X_MAX_PEND - 1; i >= 0; --i)
                if (i < np)
                    for (int k = 0; k < 4; ++k) xo[k] = xo[k] + (float)al[i] * (float)d.pp[i][k];
            for (int k = 0; k < 4; ++k) xo[k] = xo[k] + (float)af * (float)d.pv[k];
            st4(a.x32 + idx, xo);
        } else {
            double(&xd)[4] = reinterpret_cast<double(&)[4]>(d.xo);
#pragma unroll
            for (int i = X_MAX_PEND - 1; i >= 0; --i)
                if (i < np)
                    for (int k = 0; k < 4; ++k) xd[k] += (double)al[i] * (double)d.pp[i][k];
            for (int k = 0; k < 4; ++k) xd[k] += (double)af * (double)d.pv[k];
            st4(a.x64 + idx, xd);
        }
    }
    T zv[4];
    for (int k = 0; k < 4; ++k) {
        const T dm = diagm<T>(d.fw >> (8 * k));
        const T rv = fma(-af, apv[k], dm * d.zv[k]);
        zv[k] = rv / dm;
        rz += (double)rv * (double)zv[k];
        rr += (double)rv * (double)rv;
    }
    st4(zp + idx, zv);
}

// Loads as in the direction kernel: all of a pass at once, the first pass
// before the log (current direction buffer pb and pending count xmode as the
// host expects them; the log's values win when they differ).
template <typename T, bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_update(PcgArgs a, int L, int xmode, int pb) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    const Quad Q = quad_of();
    const FaceOff FO = face_off(Q);
    const int half = threadIdx.x >> 7;
    T* __restrict__ zp = Vec<T>::u(a, 0);
    const bool defer = xmode == X_DEFER;
    const bool any = P.t0 < P.t1;
    const int tf = __builtin_amdgcn_readfirstlane(any ? min(P.t0 + half, P.t1 - 1) : -1);
    // unconditional, so that the wait counts below see one path
    const FaceTiles ftf = face_row(a.nbr, tf);
    UpdIn<T, MIXED> d;
    upd_issue(d, a, pb, zp, tf, ftf, FO, defer ? -1 : xmode);
    const double paps = partial_share(a.part, P_PAP, L - 1, P.wb, P.we);
    PcgLog e = live_log(a, P.s, L - 1);
    issue_together(e, paps, 0.0, 0.0);
    if (pass_through(a, P, e, L)) return;
    const double pap_t = wg_sum(paps, s_red);
    const double alpha = fabs(pap_t) < 1e-30 ? 0.0 : e.rz / pap_t;
    const T af = (T)alpha;
    const int merge = MIXED ? e.merge : 0;
    // the pending terms (their directions are the ring buffers before this
    // update's p) and their alphas, latest first
    const int np = defer ? -1 : e.xpend;
    T al[X_MAX_PEND];
#pragma unroll
    for (int i = 0; i < X_MAX_PEND; ++i) al[i] = (T)pend_alpha(e, i);
    if (any && (e.pbuf != pb || (!defer && np != xmode)))
        upd_issue(d, a, e.pbuf, zp, tf, ftf, FO, np);
    double rz = 0.0, rr = 0.0;
    if (any) {
        int tn = __builtin_amdgcn_readfirstlane(P.t0 + 2 + half);
        FaceTiles ftn = face_row(a.nbr, tn < P.t1 ? tn : -1);
        if (P.t0 + half < P.t1)
            upd_compute(d, a, Q, af, al, np, merge, zp, (size_t)tf * GF_TILE_CELLS + FO.own, rz,
                        rr);
        for (int tb = P.t0 + 2; tb < P.t1; tb += 2) {
            const int t = tn;
            if (t >= P.t1) break;
            UpdIn<T, MIXED> n;
            upd_issue(n, a, e.pbuf, zp, t, ftn, FO, np);
            // the next pass's row is read after this pass's vector loads are out
            __builtin_amdgcn_sched_barrier(0);
            tn = __builtin_amdgcn_readfirstlane(tb + 2 + half);
            ftn = face_row(a.nbr, tn < P.t1 ? tn : -1);
            upd_compute(n, a, Q, af, al, np, merge, zp, (size_t)t * GF_TILE_CELLS + FO.own, rz,
                        rr);
        }
    }
    put_partial2(a.part, P_RZ, P_RR, L, P.b, rz, rr, s_red);
    e.pap = pap_t;
    if (defer) {  // this update's term joins the pending ones
        e.alpha3 = e.alpha2;
        e.alpha2 = e.alpha;
        e.xpend += 1;
    } else {
        e.merge = 0;
        e.xpend = 0;
    }
    e.alpha = alpha;
    e.fresh = 1;
    write_log(a, P, e, L);
}

// Chunk end of the two-kernel form: consume and decide; in mixed mode, when
// due, z = M^-1 (rhs - A (x64 + x32)) from the FP64 true residual, with
// partials of its r.r and r.z that the scene's last workgroup to arrive sums
// to finish the log (put_partial2_last).
template <bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_replace(PcgArgs a, int L) {
    __shared__ double s_red[3 * GF_WG / GF_WAVE];
    const Part P = part_of(a);
    if (P.s < 0) return;
    PcgLog e = live_log(a, P.s, L - 1);
    const double rzs = partial_share(a.part, P_RZ, L - 1, P.wb, P.we);
    const double rrs = partial_share(a.part, P_RR, L - 1, P.wb, P.we);
    issue_together(e, rzs, rrs, 0.0);
    consume_std<MIXED>(a, e, rzs, rrs, s_red);
    const bool run = MIXED && (e.upd || (e.done && !e.converged));
    if (!run) {
        write_log(a, P, e, L);
        return;
    }
    float* __restrict__ z = a.u32[0];
    double rrt = 0.0, rz = 0.0;
    FOR_QUADS(P) {
        const size_t base = (size_t)t * GF_TILE_CELLS;
        const size_t idx = base + Q.l0;
        const uint32_t fw = ld_flags(a.flags, idx);
        const FaceTiles ft = face_tiles(a.nbr, t);
        double cd[4];
        float cf[4], bv[4];
        ld4(a.x64 + idx, cd);
        ld4(a.x32 + idx, cf);
        ld4(a.b + idx, bv);
        Nb4<double> fd, od;
        Nb4<float> ff;
        load_face(a.x64, ft, FO, fd);
        load_face(a.x32, ft, FO, ff);
        double c[4];
        for (int k = 0; k < 4; ++k) {
            c[k] = cd[k] + (double)cf[k];
            fd.xm[k] += (double)ff.xm[k];
            fd.xp[k] += (double)ff.xp[k];
            fd.ym[k] += (double)ff.ym[k];
            fd.yp[k] += (double)ff.yp[k];
        }
        fd.zm0 += (double)ff.zm0;
        fd.zp3 += (double)ff.zp3;
        nb_from_lanes(c, fd, Q, od);
        float z32[4];
        for (int k = 0; k < 4; ++k) {
            const int d = fl_diag(fw, k), fl = fl_fluid(fw, k);
            const double ax = (double)d * c[k] - nb_sum4(od, c, k);
            const double rv = keep_if((double)(a.rhs_sign * bv[k]) - ax, fl);
            const float rk = (float)rv;
            const float zk = rk / (float)(d > 1 ? d : 1);
            rrt += rv * rv;
            rz += (double)rk * (double)zk;
            z32[k] = zk;
        }
        st4(z + idx, z32);
    }
    __shared__ int s_last;
    if (!put_partial2_last(a.part, P_RRT, P_RZT, L, P, rrt, rz, s_red, &s_last)) return;
    // the scene's last workgroup: the true residual and the new r.z
    // (k_pcg_fin's sums, in its order), and the chunk-end decision
    rrt = partial_share_coh(a.part, P_RRT, L, P.wb, P.we);
    rz = partial_share_coh(a.part, P_RZT, L, P.wb, P.we);
    double del = 0.0;
    wg_sum3(rrt, rz, del, s_red);
    finish_replace(a, e, rrt, rz, del);
    if (threadIdx.x == 0) {
        set_live_log(a, P.s, L, e);
        if (a.hout) put_log(a.hout + 2 * P.s + (L & 1), e);
    }
}

// pc = FP32 result (x64 [+ x32]); non-fluid cells hold 0 in both.
template <bool MIXED>
__global__ __launch_bounds__(GF_WG) void k_pcg_final(PcgArgs a) {
    const Part P = part_of(a);
    if (P.s < 0) return;
    FOR_QUADS(P) {
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
}

// Each scene's final log into the result entries (one thread per scene).
__global__ void k_pcg_result(PcgArgs a, int L) {
    const int sc = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc < a.S) put_log(a.res + sc, live_log(a, sc, L));
}

// Jacobi sweep (solver3d.py:878-938) on compact tiles: two tiles per
// workgroup, p_out = (sum of fluid-neighbour p_in - div) / diag on fluid cells
// with diag > 0, else 0 (non-fluid p_in entries are 0, so the neighbour sum in
// the reference order equals the reference's fluid-only sum).
__global__ __launch_bounds__(GF_WG) void k_jacobi(const uint8_t* __restrict__ flags,
                                                  const int32_t* __restrict__ nbr,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ pin,
                                                  float* __restrict__ pout, int nact,
                                                  const int32_t* __restrict__ nact_dev) {
    const int t = __builtin_amdgcn_readfirstlane(2 * gf::xcd_remap(blockIdx.x, gridDim.x) +
                                                 (threadIdx.x >> 7));
    if (t >= (nact_dev ? nact_dev[0] : nact)) return;
    const Quad Q = quad_of();
    const size_t base = (size_t)t * GF_TILE_CELLS;
    const size_t idx = base + Q.l0;
    const uint32_t fw = ld_flags(flags, idx);
    float c[4], bv[4], out[4];
    ld4(pin + idx, c);
    ld4(b + idx, bv);
    Nb4<float> o;
    load_nb4(pin, nb_index(nbr, t, base, Q), o);
    for (int k = 0; k < 4; ++k) {
        const int d = fl_diag(fw, k);
        const float v = (nb_sum4(o, c, k) - bv[k]) / (float)(d > 0 ? d : 1);
        out[k] = keep_if(v, fl_fluid(fw, k) & (d > 0));
    }
    st4(pout + idx, out);
}

// One red-black Gauss-Seidel half sweep (solver3d.py:1204-1267) in place on
// compact tiles: cells whose global (i+j+k) is even (color 0) or odd
// (color 1) get p = (sum of fluid-neighbour p - div) / diag, in the
// reference's neighbour order; non-fluid cells and diag 0 give 0.  Tiles
// start at multiples of 8, so the colour of quad element e is
// (li + lj + e) % 2.  A cell's six neighbours all have the other colour and
// are only read in this launch, so the in-place update has no race; the
// quad is written back whole (its other-colour cells unchanged).
__global__ __launch_bounds__(GF_WG) void k_gsrb(const uint8_t* __restrict__ flags,
                                                const int32_t* __restrict__ nbr,
                                                const float* __restrict__ b, float* p, int nact,
                                                const int32_t* __restrict__ nact_dev, int color) {
    const int t = __builtin_amdgcn_readfirstlane(2 * gf::xcd_remap(blockIdx.x, gridDim.x) +
                                                 (threadIdx.x >> 7));
    if (t >= (nact_dev ? nact_dev[0] : nact)) return;
    const Quad Q = quad_of();
    const size_t base = (size_t)t * GF_TILE_CELLS;
    const size_t idx = base + Q.l0;
    const uint32_t fw = ld_flags(flags, idx);
    float c[4], bv[4], out[4];
    ld4(p + idx, c);
    ld4(b + idx, bv);
    Nb4<float> o;
    load_nb4(p, nb_index(nbr, t, base, Q), o);
    for (int k = 0; k < 4; ++k) {
        const int d = fl_diag(fw, k);
        const float v = (nb_sum4(o, c, k) - bv[k]) / (float)(d > 0 ? d : 1);
        const float upd = keep_if(v, fl_fluid(fw, k) & (d > 0));
        out[k] = ((Q.li + Q.lj + k) & 1) == color ? upd : c[k];
    }
    st4(p + idx, out);
}

__global__ void k_fill_f32(float* a, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

namespace gf {
int scene_check(const gf_scene_t* s, const char* who);

// Solver-log readback slots in pinned host memory per host thread, grown on
// demand and kept for the life of the thread (the PCG's only host
// allocation; it never touches device memory).
// *dev (optional) receives the device-side address of the same memory, which
// kernels write directly.
static PcgLog* pinned_pcg_logs(size_t entries, PcgLog** dev = nullptr) {
    static thread_local PcgLog* logs = nullptr;
    static thread_local PcgLog* dlogs = nullptr;
    static thread_local size_t cap = 0;
    if (entries > cap) {
        if (logs) hipHostFree(logs);
        logs = dlogs = nullptr;
        cap = 0;
        void* p = nullptr;
        hipError_t e = hipHostMalloc(&p, entries * sizeof(PcgLog), hipHostMallocDefault);
        if (e != hipSuccess) {
            set_error((int)e, "hipHostMalloc: %s", hipGetErrorString(e));
            return nullptr;
        }
        void* d = nullptr;
        e = hipHostGetDevicePointer(&d, p, 0);
        if (e != hipSuccess) {
            hipHostFree(p);
            set_error((int)e, "hipHostGetDevicePointer: %s", hipGetErrorString(e));
            return nullptr;
        }
        logs = static_cast<PcgLog*>(p);
        dlogs = static_cast<PcgLog*>(d);
        cap = entries;
    }
    if (dev) *dev = dlogs;
    return logs;
}
}

extern "C" int gf_scene_jacobi(const gf_scene_t* s, int n_active, int iters, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_jacobi");
    if (rc) return rc;
    const int ntiles = s->tnx * s->tny * s->tnz;
    const bool dev_count = n_active == GF_NACT_DEVICE;
    GF_REQUIRE(dev_count || (n_active >= 0 && n_active <= ntiles), "gf_scene_jacobi: n_active %d",
               n_active);
    GF_REQUIRE(iters >= 0, "gf_scene_jacobi: iters %d", iters);
    GF_REQUIRE(s->flags && s->nbr && s->b && s->pc && s->ptmp && (s->n_active || !dev_count),
               "gf_scene_jacobi: buffers not set");
    if (n_active == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    // device count: launches sized for every tile, the kernels stop at n_active[0]
    const int span = dev_count ? ntiles : n_active;
    const int nact = dev_count ? 0 : n_active;
    const int32_t* nact_dev = dev_count ? s->n_active : nullptr;
    const size_t n = (size_t)span * GF_TILE_CELLS;
    const int fb = (int)((n + GF_WG - 1) / GF_WG);
    hipLaunchKernelGGL(k_fill_f32, dim3(fb), dim3(GF_WG), 0, st, s->pc, n, 0.0f);
    hipLaunchKernelGGL(k_fill_f32, dim3(fb), dim3(GF_WG), 0, st, s->ptmp, n, 0.0f);
    for (int k = 0; k < iters; ++k) {
        const bool to_pc = ((iters - 1 - k) % 2) == 0;
        float* out = to_pc ? s->pc : s->ptmp;
        const float* in = to_pc ? s->ptmp : s->pc;
        hipLaunchKernelGGL(k_jacobi, dim3((span + 1) / 2), dim3(GF_WG), 0, st, s->flags,
                           s->nbr, s->b, in, out, nact, nact_dev);
    }
    GF_LAUNCH_CHECK("k_jacobi");
    return 0;
}

extern "C" int gf_scene_gsrb(const gf_scene_t* s, int n_active, int iters, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_gsrb");
    if (rc) return rc;
    const int ntiles = s->tnx * s->tny * s->tnz;
    const bool dev_count = n_active == GF_NACT_DEVICE;
    GF_REQUIRE(dev_count || (n_active >= 0 && n_active <= ntiles), "gf_scene_gsrb: n_active %d",
               n_active);
    GF_REQUIRE(iters >= 0, "gf_scene_gsrb: iters %d", iters);
    GF_REQUIRE(s->flags && s->nbr && s->b && s->pc && (s->n_active || !dev_count),
               "gf_scene_gsrb: buffers not set");
    if (n_active == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int span = dev_count ? ntiles : n_active;
    const int nact = dev_count ? 0 : n_active;
    const int32_t* nact_dev = dev_count ? s->n_active : nullptr;
    const size_t n = (size_t)span * GF_TILE_CELLS;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + GF_WG - 1) / GF_WG)), dim3(GF_WG), 0, st,
                       s->pc, n, 0.0f);
    for (int k = 0; k < iters; ++k) {
        for (int color = 0; color < 2; ++color)
            hipLaunchKernelGGL(k_gsrb, dim3((span + 1) / 2), dim3(GF_WG), 0, st, s->flags,
                               s->nbr, s->b, s->pc, nact, nact_dev, color);
    }
    GF_LAUNCH_CHECK("k_gsrb");
    return 0;
}

namespace {

// Validated solver arguments shared by gf_scene_pcg_sync and gf_scene_pcg_enqueue.
// With own_work = false the work buffers (vectors, partials, live log,
// partition table) are left for the caller to point at its own memory, and
// only the descriptor's inputs, pc and the result entries of scalars are
// required.
// The two-kernel form's packed flags live in its unused w~ buffer.
void set_pflags(PcgArgs& a, bool mixed) {
    a.pflags = a.variant == PCG_STD
                   ? reinterpret_cast<uint16_t*>(mixed ? static_cast<void*>(a.w32[0])
                                                       : static_cast<void*>(a.w64[0]))
                   : nullptr;
}

int pcg_prepare(const gf_scene_t* s, int n_active, int precision, double tol, int max_iter,
                int rhs_sign, const char* who, PcgArgs& a, bool own_work = true) {
    int rc = gf::scene_check(s, who);
    if (rc) return rc;
    const int ntiles = s->tnx * s->tny * s->tnz;
    const bool packed = s->n_scenes > 1;
    const int S = packed ? s->n_scenes : 1;
    // GF_NACT_DEVICE: the active tile count is read from n_active[0] on the
    // device, so the launches do not depend on it (capture-safe); grid and
    // form are then sized for every tile of the grid
    const bool dev_count = n_active == GF_NACT_DEVICE && !packed;
    GF_REQUIRE(dev_count || (n_active >= 0 && n_active <= ntiles), "%s: n_active %d", who,
               n_active);
    GF_REQUIRE(!dev_count || s->n_active, "%s: n_active buffer not set", who);
    const int sized = dev_count ? ntiles : n_active;
    GF_REQUIRE(S <= PCG_GMAX, "%s: %d scenes exceed %d", who, S, PCG_GMAX);
    GF_REQUIRE(precision == GF_PCG_MIXED || precision == GF_PCG_FP64, "%s: precision %d", who,
               precision);
    GF_REQUIRE(rhs_sign == 1 || rhs_sign == -1, "%s: rhs_sign %d", who, rhs_sign);
    GF_REQUIRE(tol >= 0.0 && max_iter >= 0, "%s: tol/max_iter", who);
    GF_REQUIRE(s->flags && s->nbr && s->b && s->pc && s->scalars, "%s: buffers not set", who);
    if (own_work) {
        GF_REQUIRE(s->partials && s->x64 && s->pcg_part, "%s: buffers not set", who);
        // FP32 vectors serve both precisions' replacement bookkeeping only in mixed mode
        if (precision == GF_PCG_MIXED)
            GF_REQUIRE(s->x32 && s->r32 && s->p32 && s->p32b && s->ap32 && s->cg32,
                       "%s: FP32 vectors not set", who);
        else
            GF_REQUIRE(s->r64 && s->p64 && s->p64b && s->ap64 && s->cg64,
                       "%s: FP64 vectors not set", who);
    }
    const size_t C = (size_t)ntiles * GF_TILE_CELLS;
    a.flags = s->flags; a.nbr = s->nbr; a.b = s->b;
    a.u32[0] = s->r32; a.u32[1] = s->cg32;
    a.w32[0] = s->ap32; a.w32[1] = s->cg32 ? s->cg32 + C : nullptr;
    a.s32[0] = s->p32b; a.s32[1] = s->cg32 ? s->cg32 + 2 * C : nullptr;
    a.p32 = s->p32; a.x32 = s->x32;
    a.u64[0] = s->r64; a.u64[1] = s->cg64;
    a.w64[0] = s->ap64; a.w64[1] = s->cg64 ? s->cg64 + C : nullptr;
    a.s64[0] = s->p64b; a.s64[1] = s->cg64 ? s->cg64 + 2 * C : nullptr;
    a.p64 = s->p64; a.x64 = s->x64;
    a.pc = s->pc; a.part = s->partials; a.log = (PcgLog*)s->scalars; a.hout = nullptr;
    a.res = a.log + 2 * S;
    a.part_info = s->pcg_part;
    a.scene_slot = packed ? s->scene_slot : nullptr;
    a.nact = dev_count ? 0 : n_active;
    a.nact_dev = dev_count ? s->n_active : nullptr;
    a.S = S;
    // 1024 workgroups (4 per CU) measured best on MI355X at 256^3 (5.4 M
    // active cells); the autotuner may choose another value up to PCG_GMAX.
    int gmax = gf::tunable(gf::GF_TUNE_PCG_WORKGROUPS);
    gmax = gmax > 0 ? (gmax < PCG_GMAX ? gmax : PCG_GMAX) : PCG_DEFAULT_G;
    if (packed) {
        a.G = gmax > S ? gmax : S;
    } else {
        const int pairs = (sized + 1) / 2;
        a.G = pairs < gmax ? pairs : gmax;
    }
    const int vt = gf::tunable(gf::GF_TUNE_PCG_VARIANT);
    if (vt == PCG_STD || vt == PCG_CG1)
        a.variant = vt;
    else
        a.variant = sized < PCG_CG1_MAX_TILES ? PCG_CG1 : PCG_STD;
    a.rhs_sign = (float)rhs_sign;
    a.tol2 = tol * tol;
    // replace when r.r has fallen 1000x since the last replacement.  The
    // iteration count to a given residual does not depend on this interval
    // (measured at 256^3 and 128^3: 4 to 56 replacements per solve, the same
    // count to 1e-8), so the longest interval that still catches drift is
    // kept: each replacement is one extra launch of about 36 us at 256^3.
    a.delta2 = 1e-3;
    a.max_iter = max_iter;
    set_pflags(a, precision == GF_PCG_MIXED);
    return 0;
}

// Kernels per iteration and launch numbers used by one chunk of `chunk` iterations.
int pcg_kernels_per_iter(const PcgArgs& a) { return a.variant == PCG_CG1 ? 1 : 2; }
int pcg_chunk_launches(const PcgArgs& a, int chunk, bool mixed) {
    return a.variant == PCG_CG1 ? chunk + 1 + (mixed ? 1 : 0) : 2 * chunk + 1;
}

void pcg_launch_setup(const PcgArgs& a, bool mixed, hipStream_t st) {
    const dim3 grid(a.G), block(GF_WG);
    hipLaunchKernelGGL(k_pcg_partition, dim3(1), block, 0, st, a);
    if (a.variant == PCG_STD) {
        if (mixed) hipLaunchKernelGGL((k_pcg_setup<float, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_pcg_setup<double, false>), grid, block, 0, st, a);
    } else if (mixed) {
        hipLaunchKernelGGL((k_cg1_setup<float, true>), grid, block, 0, st, a);
        hipLaunchKernelGGL(k_cg1_setup2<float>, grid, block, 0, st, a);
    } else {
        hipLaunchKernelGGL((k_cg1_setup<double, false>), grid, block, 0, st, a);
        hipLaunchKernelGGL(k_cg1_setup2<double>, grid, block, 0, st, a);
    }
    hipLaunchKernelGGL(k_pcg_fin, dim3(a.S), block, 0, st, a, 1);
}

// One chunk: `chunk` iterations, then the chunk-end decision (and residual
// replacement when due).  L is the launch number before the chunk.  ev, when
// given, holds 2*k+2 events (k = pcg_kernels_per_iter): the start and stop of
// each kernel of the chunk's iteration jsample, then of the chunk-end
// replacement, taken from the kernels' own dispatch packets
// (hipExtLaunchKernel), so they time the kernels' execution and not the gaps
// between launches.  One iteration per chunk is a sample that keeps the
// timing's own cost out of the other launches; the caller varies jsample
// over chunks (the update kernel adds x's pending terms on some iterations
// only).
int pcg_launch_chunk(const PcgArgs& a, bool mixed, int chunk, int& L, hipStream_t st,
                     hipEvent_t* ev, PcgLog* hout = nullptr, int jsample = -1) {
    if (jsample < 0) jsample = chunk / 2;
    const dim3 grid(a.G), block(GF_WG);
    const int kpi = pcg_kernels_per_iter(a);
    // the chunk's last kernel also writes the logs to hout (host readback
    // without a copy on the stream)
    PcgArgs al = a;
    al.hout = hout;
    auto e0 = [&](int i) { return ev ? ev[2 * i] : nullptr; };
    auto e1 = [&](int i) { return ev ? ev[2 * i + 1] : nullptr; };
    hipEvent_t* const ev_all = ev;
    // timed launches carry their events on the dispatch packet; untimed ones
    // (also those recorded into a graph) are plain launches
#define GF_TLAUNCH(K, E0, E1, ...)                                                  \
    do {                                                                            \
        if (ev) hipExtLaunchKernelGGL(K, grid, block, 0, st, E0, E1, 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(K, grid, block, 0, st, __VA_ARGS__);                \
    } while (0)
    // two-kernel form: x on every xe-th update and on the chunk's last one
    // (tunable pcg_x_every, 1..4; default 4)
    int xe = gf::tunable(gf::GF_TUNE_PCG_X_EVERY);
    xe = xe >= 1 && xe <= X_MAX_PEND + 1 ? xe : X_MAX_PEND + 1;
    const int c0 = (L - 1) / pcg_chunk_launches(a, chunk, mixed);  // chunks before this one
    for (int j = 0; j < chunk; ++j) {
        const int i = 0;
        ev = j == jsample ? ev_all : nullptr;
        const bool xadd = j % xe == xe - 1 || j + 1 == chunk;
        const int xmode = xadd ? j % xe : X_DEFER;
        // direction buffer of the k-th direction kernel since the setup
        // (k = 1, 2, ...): p_old in ring slot (k-1)&3, p_new in k&3, for every
        // scene still iterating (a stopped scene passes through)
        const int k = c0 * chunk + j + 1;
        if (a.variant == PCG_CG1) {
            ++L;
            if (mixed)
                GF_TLAUNCH((k_cg1_iter<float, true>), e0(i), e1(i), a, L);
            else
                GF_TLAUNCH((k_cg1_iter<double, false>), e0(i), e1(i), a, L);
        } else {
            ++L;
            if (mixed)
                GF_TLAUNCH((k_pcg_dirspmv<float, true>), e0(i), e1(i), a, L, (k - 1) & 3);
            else
                GF_TLAUNCH((k_pcg_dirspmv<double, false>), e0(i), e1(i), a, L, (k - 1) & 3);
            ++L;
            if (mixed)
                GF_TLAUNCH((k_pcg_update<float, true>), e0(i + 1), e1(i + 1), a, L, xmode, k & 3);
            else
                GF_TLAUNCH((k_pcg_update<double, false>), e0(i + 1), e1(i + 1), a, L, xmode,
                           k & 3);
        }
    }
    const int ic = kpi;
    ev = ev_all;
    ++L;
    if (a.variant == PCG_CG1) {
        if (mixed) {
            GF_TLAUNCH(k_cg1_replace<true>, e0(ic), nullptr, a, L);
            ++L;
            GF_TLAUNCH(k_cg1_replace2, nullptr, e1(ic), al, L);
        } else {
            GF_TLAUNCH(k_cg1_replace<false>, e0(ic), e1(ic), al, L);
        }
    } else {
        if (mixed)
            GF_TLAUNCH(k_pcg_replace<true>, e0(ic), e1(ic), al, L);
        else
            GF_TLAUNCH(k_pcg_replace<false>, e0(ic), e1(ic), al, L);
    }
    // a mixed replacement finishes in its last workgroup (no k_pcg_fin)
#undef GF_TLAUNCH
    return gf::check_launch("pcg iteration");
}

void pcg_launch_final(const PcgArgs& a, bool mixed, int L, hipStream_t st) {
    const dim3 grid(a.G), block(GF_WG);
    if (mixed)
        hipLaunchKernelGGL(k_pcg_final<true>, grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(k_pcg_final<false>, grid, block, 0, st, a);
    hipLaunchKernelGGL(k_pcg_result, dim3((a.S + GF_WG - 1) / GF_WG), block, 0, st, a, L);
}

}  // namespace

extern "C" int gf_scene_pcg_sync(const gf_scene_t* s, int n_active, int precision, double tol,
                                 int max_iter, int check_every, int rhs_sign, int* iters,
                                 double* rel_residual, int* replacements, double* kernel_ms,
                                 void* stream) {
    if (iters) *iters = 0;
    if (rel_residual) *rel_residual = 0.0;
    if (replacements) *replacements = 0;
    if (kernel_ms) for (int i = 0; i < 4; ++i) kernel_ms[i] = 0.0;
    PcgArgs a = {};
    GF_REQUIRE(n_active != GF_NACT_DEVICE, "gf_scene_pcg_sync: needs the host tile count");
    int rc = pcg_prepare(s, n_active, precision, tol, max_iter, rhs_sign, "gf_scene_pcg_sync", a);
    if (rc) return rc;
    if (check_every < 1) check_every = 1;
    if (check_every > 64) check_every = 64;
    hipStream_t st = (hipStream_t)stream;
    const int S = a.S;
    if (S == 1 && n_active == 0) return 0;
    const bool mixed = precision == GF_PCG_MIXED;
    pcg_launch_setup(a, mixed, st);
    int L = 1;
    GF_LAUNCH_CHECK("k_pcg_partition/setup/fin");

    // The host reads the scene logs back once per chunk of check_every
    // iterations.  Chunks are pipelined: chunk c+1 is enqueued before the host
    // waits for chunk c's logs (copied into pinned memory), so the device does
    // not idle during the round trip; once every scene has stopped, the
    // kernels of the chunk already in flight pass through.
    // Optional per-kernel timing of the iterations that did work: events
    // around every launch of a chunk, read when the chunk's logs are read.
    // kernel_ms: [0] iteration kernel (two-kernel form: direction+SpMV), [1] update
    // (two-kernel form), summed over the sampled iterations (one per chunk,
    // when it did work), [2] the number of sampled iterations, [3]
    // chunk-end replacements (every one that ran).
    PcgLog* dlog = nullptr;  // device-side address of hlog
    PcgLog* hlog = gf::pinned_pcg_logs((size_t)6 * S, &dlog);  // setup + 2 chunk slots, 2S entries each
    if (!hlog) return gf::set_error(GF_ESTATE, "gf_scene_pcg_sync: pinned log: %s", gf_last_error());
    const bool timed = kernel_ms != nullptr;
    const int chunk = check_every;
    const int kpi = pcg_kernels_per_iter(a);
    const int nev = timed ? 2 * kpi + 2 : 0;
    hipEvent_t ev[2][2 * 2 + 2];
    hipEvent_t done_ev[2] = {nullptr, nullptr};
    int created = 0;
    hipError_t e = hipSuccess;
    for (int sl = 0; sl < 2 && e == hipSuccess; ++sl) {
        e = hipEventCreateWithFlags(&done_ev[sl], hipEventDisableTiming);
        for (int i = 0; i < nev && e == hipSuccess; ++i) {
            e = hipEventCreate(&ev[sl][i]);
            if (e == hipSuccess) created = sl * nev + i + 1;
        }
    }
    auto destroy_events = [&]() {
        for (int sl = 0; sl < 2; ++sl) {
            if (done_ev[sl]) hipEventDestroy(done_ev[sl]);
            for (int i = 0; i < nev && sl * nev + i < created; ++i) hipEventDestroy(ev[sl][i]);
        }
    };
    if (e != hipSuccess) {
        destroy_events();
        return gf::set_error((int)e, "gf_scene_pcg_sync: event: %s", hipGetErrorString(e));
    }
    // Aggregate of the scene logs in one slot read at launch parity L.
    struct Agg {
        int done, iters, nupd;
        double worst;  // max relative true residual
    };
    auto aggregate = [&](const PcgLog* lg, int Lp) {
        Agg g = {1, 0, 0, 0.0};
        for (int sc = 0; sc < S; ++sc) {
            const PcgLog& q = lg[2 * sc + (Lp & 1)];
            g.done &= q.done;
            g.iters = q.iters > g.iters ? q.iters : g.iters;
            g.nupd += q.nupd;
            const double r = q.bb > 0.0 ? sqrt(q.rr_true / q.bb) : 0.0;
            g.worst = r > g.worst ? r : g.worst;
        }
        return g;
    };
    const size_t live_bytes = (size_t)2 * S * sizeof(PcgLog);
    Agg hs = {0, 0, 0, 0.0};
    e = hipMemcpyAsync(hlog, a.log, live_bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = gf::set_error((int)e, "gf_scene_pcg_sync: readback: %s", hipGetErrorString(e));
    else hs = aggregate(hlog, L);

    // the timed iteration of a chunk: its position cycles through the chunk
    // over consecutive timed chunks (of all solves in the process), so that
    // the iterations that add x's pending terms are sampled as often as
    // they occur
    static std::atomic<unsigned> sample_rot{0};
    int js[2] = {0, 0};
    int launched = 0, next_slot = 0, npend = 0, pend[2] = {0, 0}, pend_L[2] = {0, 0};
    auto enqueue_chunk = [&]() -> int {
        const int sl = next_slot;
        js[sl] = timed ? (int)(sample_rot.fetch_add(1, std::memory_order_relaxed) % (unsigned)chunk)
                       : 0;
        int r = pcg_launch_chunk(a, mixed, chunk, L, st, timed ? ev[sl] : nullptr,
                                 dlog + (size_t)2 * S * (1 + sl), js[sl]);
        if (r) return r;
        hipError_t ce = hipEventRecord(done_ev[sl], st);
        if (ce != hipSuccess)
            return gf::set_error((int)ce, "gf_scene_pcg_sync: readback: %s", hipGetErrorString(ce));
        pend[npend] = sl;
        pend_L[npend] = L;
        ++npend;
        next_slot ^= 1;
        launched += chunk;
        return 0;
    };
    const int launch_cap = max_iter + 8 * chunk;
    int iters_before = hs.iters, nupd_before = hs.nupd;
    if (!rc && !hs.done) rc = enqueue_chunk();
    while (!rc && npend > 0) {
        if (npend < 2 && launched < launch_cap) {
            rc = enqueue_chunk();  // look-ahead
            if (rc) break;
        }
        const int sl = pend[0], Lsl = pend_L[0];
        pend[0] = pend[1];
        pend_L[0] = pend_L[1];
        --npend;
        e = hipEventSynchronize(done_ev[sl]);
        if (e != hipSuccess) {
            rc = gf::set_error((int)e, "gf_scene_pcg_sync: readback: %s", hipGetErrorString(e));
            break;
        }
        hs = aggregate(hlog + (size_t)2 * S * (1 + sl), Lsl);
        if (timed) {
            const int real = hs.iters - iters_before;
            if (real > js[sl]) {  // the sampled iteration did work
                for (int m = 0; m < kpi; ++m) {
                    float t0 = 0.f;
                    (void)hipEventElapsedTime(&t0, ev[sl][2 * m], ev[sl][2 * m + 1]);
                    kernel_ms[m] += t0;
                }
                kernel_ms[2] += 1.0;
            }
            if (mixed && hs.nupd > nupd_before) {
                const int i = kpi;
                float t3 = 0.f;
                (void)hipEventElapsedTime(&t3, ev[sl][2 * i], ev[sl][2 * i + 1]);
                kernel_ms[3] += t3;
            }
        }
        iters_before = hs.iters;
        nupd_before = hs.nupd;
        if (hs.done) break;
    }
    // let the chunk still in flight (passing through) finish before its
    // events and its pinned log slot are released
    for (int q = 0; q < npend; ++q) {
        hipError_t we = hipEventSynchronize(done_ev[pend[q]]);
        if (we != hipSuccess && !rc)
            rc = gf::set_error((int)we, "gf_scene_pcg_sync: drain: %s", hipGetErrorString(we));
    }
    destroy_events();
    if (rc) return rc;
    pcg_launch_final(a, mixed, L, st);
    GF_LAUNCH_CHECK("k_pcg_final");
    if (iters) *iters = hs.iters;
    if (rel_residual) *rel_residual = hs.worst;
    if (replacements) *replacements = hs.nupd;
    return 0;
}

extern "C" int gf_scene_pcg_enqueue(const gf_scene_t* s, int n_active, int precision, double tol,
                                    int max_iter, int check_every, int rhs_sign, int first_chunk,
                                    int n_chunks, void* stream) {
    PcgArgs a = {};
    int rc = pcg_prepare(s, n_active, precision, tol, max_iter, rhs_sign, "gf_scene_pcg_enqueue", a);
    if (rc) return rc;
    GF_REQUIRE(check_every >= 1 && check_every <= 64, "gf_scene_pcg_enqueue: check_every %d",
               check_every);
    GF_REQUIRE(first_chunk >= 0 && n_chunks >= 0, "gf_scene_pcg_enqueue: chunks %d+%d", first_chunk,
               n_chunks);
    if (a.S == 1 && n_active == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const bool mixed = precision == GF_PCG_MIXED;
    if (first_chunk == 0) pcg_launch_setup(a, mixed, st);
    int L = 1 + first_chunk * pcg_chunk_launches(a, check_every, mixed);
    for (int c = 0; c < n_chunks; ++c) {
        rc = pcg_launch_chunk(a, mixed, check_every, L, st, nullptr);
        if (rc) return rc;
    }
    pcg_launch_final(a, mixed, L, st);
    GF_LAUNCH_CHECK("gf_scene_pcg_enqueue");
    return 0;
}

extern "C" int gf_scene_pcg_stats(const gf_scene_t* s, int* iters, double* rel_residual,
                                  int* replacements, void* stream) {
    int rc = gf::scene_check(s, "gf_scene_pcg_stats");
    if (rc) return rc;
    GF_REQUIRE(s->scalars, "gf_scene_pcg_stats: scalars not set");
    const int S = s->n_scenes > 1 ? s->n_scenes : 1;
    PcgLog* hlog = gf::pinned_pcg_logs((size_t)S);
    if (!hlog) return gf::set_error(GF_ESTATE, "gf_scene_pcg_stats: pinned log: %s", gf_last_error());
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(hlog, (const PcgLog*)s->scalars + 2 * S, S * sizeof(PcgLog),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return gf::set_error((int)e, "gf_scene_pcg_stats: readback: %s", hipGetErrorString(e));
    for (int sc = 0; sc < S; ++sc) {
        if (iters) iters[sc] = hlog[sc].iters;
        if (rel_residual) rel_residual[sc] = hlog[sc].bb > 0.0 ? sqrt(hlog[sc].rr_true / hlog[sc].bb) : 0.0;
        if (replacements) replacements[sc] = hlog[sc].nupd;
    }
    return 0;
}

// The reference's on-device stop flag and threshold of its last PCG solve
// (k3_zero_int_scalar / k3_set_tol_sq / k3_check_converged,
// solver3d.py:1916-1942, 2949-2983): done = 1 when the solve met its
// tolerance, tol_sq = tol^2 * ||r0||^2 in FP32 (r0 = b).  A zero right-hand
// side leaves done = 0, as the reference's early return does.  One thread
// reads scene 0's result log.
__global__ void k_pcg_flags(const PcgLog* __restrict__ res, float tol, int32_t* done,
                            float* tol_sq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (done) done[0] = (res->converged && res->bb > 0.0) ? 1 : 0;
    if (tol_sq) tol_sq[0] = tol * tol * (float)res->bb;
}

extern "C" int gf_scene_pcg_flags(const gf_scene_t* s, float tol, int32_t* done, float* tol_sq,
                                  void* stream) {
    int rc = gf::scene_check(s, "gf_scene_pcg_flags");
    if (rc) return rc;
    GF_REQUIRE(s->scalars, "gf_scene_pcg_flags: scalars not set");
    GF_REQUIRE(done || tol_sq, "gf_scene_pcg_flags: no output");
    const int S = s->n_scenes > 1 ? s->n_scenes : 1;
    const PcgLog* res = reinterpret_cast<const PcgLog*>(s->scalars) + 2 * S;
    hipLaunchKernelGGL(k_pcg_flags, dim3(1), dim3(64), 0, (hipStream_t)stream, res, tol, done,
                       tol_sq);
    GF_LAUNCH_CHECK("gf_scene_pcg_flags");
    return 0;
}

// ---- no-allocation PCG ---------------------------------------------------
// gf_scene_pcg_enqueue_noalloc runs the same solve with every work buffer in
// one file-scope device array, so a caller that cannot allocate (code under
// hipGraph capture, serving loops) passes only the inputs and outputs.
// Layout (bytes): the FP64 iterate x64 (8 per slot), then seven work vectors
// of 8 bytes per slot (FP64 mode: u, w~, s~ ping-pong pairs and p; mixed mode
// uses the first half of each as FP32, plus x32), the per-workgroup partials,
// the live log pairs and the workgroup partition table.  Every word a call
// reads it has written earlier in the same call (the setup kernels write all
// 512 entries of every active slot), so nothing needs clearing; calls that
// overlap on one GPU share the array and race.
namespace {
constexpr int NOALLOC_MAX_TILES = GF_PCG_NOALLOC_MAX_TILES;
constexpr int NOALLOC_MAX_SCENES = GF_PCG_NOALLOC_MAX_SCENES;
constexpr size_t NOALLOC_SLOTS = (size_t)NOALLOC_MAX_TILES * GF_TILE_CELLS;
constexpr size_t NOALLOC_VEC_BYTES = 8 * 8 * NOALLOC_SLOTS;                     // 64 MiB
constexpr size_t NOALLOC_PART_BYTES = sizeof(double) * 2 * P_SET * PCG_GMAX;     // 256 KiB
constexpr size_t NOALLOC_LOG_BYTES = sizeof(PcgLog) * 2 * NOALLOC_MAX_SCENES;    // 16 KiB
constexpr size_t NOALLOC_INFO_BYTES =
    sizeof(int32_t) * ((8 * PCG_GMAX + NOALLOC_MAX_SCENES + 1 + 63) / 64 * 64);
constexpr size_t NOALLOC_BYTES =
    NOALLOC_VEC_BYTES + NOALLOC_PART_BYTES + NOALLOC_LOG_BYTES + NOALLOC_INFO_BYTES;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static_assert(NOALLOC_BYTES == GF_PCG_NOALLOC_SCRATCH_BYTES, "header states the scratch size");

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name of a __device__ array is not its address).
int scratch_address(void** out, const char* who) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return gf::set_error((int)e, "%s: hipGetDevice: %s", who, hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return gf::set_error(GF_ERANGE, "%s: device %d", who, dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gf::set_error(e != hipSuccess ? (int)e : GF_ESTATE,
                                 "%s: hipGetSymbolAddress(scratch): %s", who,
                                 hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" int gf_scene_pcg_enqueue_noalloc(const gf_scene_t* s, int n_active, int precision,
                                            double tol, int max_iter, int check_every,
                                            int rhs_sign, int first_chunk, int n_chunks,
                                            void* stream) {
    const char* who = "gf_scene_pcg_enqueue_noalloc";
    PcgArgs a = {};
    int rc = pcg_prepare(s, n_active, precision, tol, max_iter, rhs_sign, who, a, false);
    if (rc) return rc;
    GF_REQUIRE(check_every >= 1 && check_every <= 64, "%s: check_every %d", who, check_every);
    GF_REQUIRE(first_chunk >= 0 && n_chunks >= 0, "%s: chunks %d+%d", who, first_chunk, n_chunks);
    const int ntiles = s->tnx * s->tny * s->tnz;
    // slots the kernels may touch: the host count, or every tile of the grid
    // when the count is read on the device
    const int reach = n_active == GF_NACT_DEVICE ? ntiles : n_active;
    if (reach > NOALLOC_MAX_TILES)
        return gf::set_error(GF_ERANGE,
                             "%s: %d %s tiles exceed the %d of the static work array", who, reach,
                             n_active == GF_NACT_DEVICE ? "grid" : "active", NOALLOC_MAX_TILES);
    if (a.S > NOALLOC_MAX_SCENES)
        return gf::set_error(GF_ERANGE, "%s: %d scenes exceed %d", who, a.S, NOALLOC_MAX_SCENES);
    if (a.S == 1 && n_active == 0) return 0;
    void* base = nullptr;
    rc = scratch_address(&base, who);
    if (rc) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    const size_t V = 8 * NOALLOC_SLOTS;   // bytes per 8-byte vector
    double* x64 = reinterpret_cast<double*>(b);
    unsigned char* vec = b + V;           // seven vectors of V bytes
    a.x64 = x64;
    if (precision == GF_PCG_MIXED) {
        float* f[8];
        for (int i = 0; i < 7; ++i) f[i] = reinterpret_cast<float*>(vec + i * V);
        f[7] = reinterpret_cast<float*>(vec + 6 * V) + NOALLOC_SLOTS;  // x32: second half of the 7th
        a.u32[0] = f[0]; a.u32[1] = f[1]; a.w32[0] = f[2]; a.w32[1] = f[3];
        a.s32[0] = f[4]; a.s32[1] = f[5]; a.p32 = f[6]; a.x32 = f[7];
    } else {
        double* d[7];
        for (int i = 0; i < 7; ++i) d[i] = reinterpret_cast<double*>(vec + i * V);
        a.u64[0] = d[0]; a.u64[1] = d[1]; a.w64[0] = d[2]; a.w64[1] = d[3];
        a.s64[0] = d[4]; a.s64[1] = d[5]; a.p64 = d[6];
    }
    set_pflags(a, precision == GF_PCG_MIXED);
    unsigned char* tail = b + NOALLOC_VEC_BYTES;
    a.part = reinterpret_cast<double*>(tail);
    a.log = reinterpret_cast<PcgLog*>(tail + NOALLOC_PART_BYTES);
    a.part_info = reinterpret_cast<int32_t*>(tail + NOALLOC_PART_BYTES + NOALLOC_LOG_BYTES);
    a.res = reinterpret_cast<PcgLog*>(s->scalars) + 2 * a.S;   // results: the caller's scalars
    hipStream_t st = (hipStream_t)stream;
    const bool mixed = precision == GF_PCG_MIXED;
    if (first_chunk == 0) pcg_launch_setup(a, mixed, st);
    int L = 1 + first_chunk * pcg_chunk_launches(a, check_every, mixed);
    for (int c = 0; c < n_chunks; ++c) {
        rc = pcg_launch_chunk(a, mixed, check_every, L, st, nullptr);
        if (rc) return rc;
    }
    pcg_launch_final(a, mixed, L, st);
    GF_LAUNCH_CHECK(who);
    return 0;
}
