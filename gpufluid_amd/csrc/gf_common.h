// Shared device helpers for the gfx950 FLIP substep kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/gpufluid_hip.h"

#define GF_WAVE 64
#define GF_WG 256

namespace gf {

// Host-side error recording (gf_api.hip).
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);

// Launch-shape tunables set through gf_tune_set (0 = built-in choice).
enum { GF_TUNE_PCG_WORKGROUPS = 0, GF_TUNE_PCG_VARIANT, GF_TUNE_PCG_X_EVERY, GF_TUNE_COUNT };
int tunable(int id);

// Exclusive scan of n int32 (gf_sort.hip): out[0..n] with out[n] = total;
// part holds ceil(n/4096)+1 entries.
int scan_exclusive_i32(const int32_t* in, int n, int32_t* out, int32_t* part, hipStream_t st);

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of work items so that neighbouring tiles (which share halo
// data) land in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int nx = 8;
    if (nblocks < nx * 2) return bid;
    int xcd = bid % nx, local = bid / nx;
    int per = nblocks / nx, rem = nblocks % nx;
    int start = xcd * per + (xcd < rem ? xcd : rem);
    return start + local;
}

// Cell binning of solver3d.py:96-99: clamp(int(p/dx), 0, n-1) with an IEEE
// FP32 divide and truncation.  The quotient is clamped to [-1, n] first so the
// conversion is always in range (NaN maps to cell 0 as the saturating
// conversion does); in-range quotients are unaffected.
__device__ __forceinline__ int bin_axis(float p, float dx, int n) {
    float q = p / dx;
    q = fminf(fmaxf(q, -1.0f), (float)n);
    int c = (int)q;
    return clampi(c, 0, n - 1);
}

// Packed scenes: per-scene table of GF_SCENE_PAR floats (include/gpufluid_hip.h,
// gf_scene_t.scene_par): [0] g*dt, [1..3] clamp lo, [4..6] clamp hi (world
// units), [7] the scene's dx, [8..10] its cell offset in the packed grid,
// [11] its cells per axis.
constexpr int SCENE_PAR = GF_SCENE_PAR;

// The grid a tile's particles live in: their own scene's dx, the scene's
// cell offset in the packed grid and its size (one scene: the whole grid).
// Positions stay in the scene's world units, so every per-particle quantity
// (cell, weights, samples) is computed exactly as for the scene run alone
// and only integer indices are shifted by the offset.
struct Frame {
    float dx;
    int ox, oy, oz, nx, ny, nz;
};

__device__ __forceinline__ Frame tile_frame(const int32_t* tile_scene, const float* scene_par,
                                            int t, float dx, int nx, int ny, int nz) {
    Frame F;
    if (!tile_scene) {
        F.dx = dx;
        F.ox = F.oy = F.oz = 0;
        F.nx = nx; F.ny = ny; F.nz = nz;
        return F;
    }
    // t is uniform over the workgroup: readfirstlane keeps the frame in
    // scalar registers (otherwise it costs the P2G kernels a wave of occupancy)
    const int sc = __builtin_amdgcn_readfirstlane(tile_scene[t]);
    const float* sp = scene_par + SCENE_PAR * (sc > 0 ? sc : 0);
    F.dx = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(sp[7])));
    F.ox = __builtin_amdgcn_readfirstlane((int)sp[8]);
    F.oy = __builtin_amdgcn_readfirstlane((int)sp[9]);
    F.oz = __builtin_amdgcn_readfirstlane((int)sp[10]);
    F.nx = F.ny = F.nz = __builtin_amdgcn_readfirstlane((int)sp[11]);
    return F;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Workgroup-wide exclusive scan of one int per thread (blockDim.x <= 1024,
// multiple of 64). Returns the exclusive prefix; *total gets the sum.
__device__ __forceinline__ int wg_exclusive_scan(int v, int* s_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < nwaves; ++i) { int t = s_wave[i]; s_wave[i] = acc; acc += t; }
        s_wave[nwaves] = acc;
    }
    __syncthreads();
    int res = s_wave[wave] + incl - v;
    *total = s_wave[nwaves];
    __syncthreads();
    return res;
}

}  // namespace gf

// Launch helpers (host).
#define GF_LAUNCH_CHECK(what)                          \
    do {                                               \
        int _rc = gf::check_launch(what);              \
        if (_rc) return _rc;                           \
    } while (0)

#define GF_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) return gf::set_error(GF_EINVAL, __VA_ARGS__); \
    } while (0)
