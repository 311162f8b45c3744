// Analytic animated obstacles for the per-frame hook (D4.6 in prepare_frame,
// solver3d.py:3918-3940, 2537-2596; SDFs of primitives/sdf.py:40-75).  One
// pass over the cells rebuilds the marker (base marker, then every obstacle's
// inside cells solid) and one pass over the faces writes the moving-boundary
// velocities: a face takes the velocity of the last obstacle, in list order,
// that contains one of its two cells (the reference's per-obstacle launches
// overwrite in that order), else 0.  The SDFs are evaluated at the FP32 cell
// centres ((i + 0.5) * dx rounded once) with each operation rounded on its
// own, and only their sign is used.
#include "gf_common.h"

namespace {

enum { OBS_SPHERE = 0, OBS_BOX = 1, OBS_CYLINDER_Y = 2 };
constexpr int OBS_MAX = 64;
constexpr int OBS_STRIDE = 10;  // kind, cx, cy, cz, p0, p1, p2, vx, vy, vz

struct ObsArgs {
    const float* obs;  // n_obs * OBS_STRIDE
    int n_obs, nx, ny, nz;
    double dx;
};

__device__ __forceinline__ float centre(int i, double dx) {
    return (float)(((double)i + 0.5) * dx);
}

// sdf <= 0 for the obstacle at cell (i, j, k)
__device__ bool inside(const float* o, float x, float y, float z) {
#pragma clang fp contract(off)
    const int kind = (int)o[0];
    const float lx = x - o[1], ly = y - o[2], lz = z - o[3];
    if (kind == OBS_SPHERE) {
        // norm(g - c) - radius (sdf_sphere, sdf.py:40-43)
        const float d = sqrtf(lx * lx + ly * ly + lz * lz) - o[4];
        return d <= 0.0f;
    }
    if (kind == OBS_BOX) {
        // every |g - c| - h <= 0 (sdf_box, sdf.py:46-53)
        return fabsf(lx) - o[4] <= 0.0f && fabsf(ly) - o[5] <= 0.0f && fabsf(lz) - o[6] <= 0.0f;
    }
    // radial and axial excess both <= 0 (sdf_cylinder_y, sdf.py:56-66)
    const float radial = sqrtf(lx * lx + lz * lz);
    return radial - o[4] <= 0.0f && fabsf(ly) - o[5] <= 0.0f;
}

// Last obstacle (index) containing cell (i, j, k), or -1.
__device__ int last_inside(const ObsArgs& a, const float* s_obs, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= a.nx || j >= a.ny || k >= a.nz) return -1;
    const float x = centre(i, a.dx), y = centre(j, a.dx), z = centre(k, a.dx);
    for (int b = a.n_obs - 1; b >= 0; --b)
        if (inside(s_obs + OBS_STRIDE * b, x, y, z)) return b;
    return -1;
}

__device__ __forceinline__ void stage(const ObsArgs& a, float* s_obs) {
    for (int e = threadIdx.x; e < a.n_obs * OBS_STRIDE; e += blockDim.x) s_obs[e] = a.obs[e];
    __syncthreads();
}

__global__ __launch_bounds__(GF_WG) void k_obs_marker(ObsArgs a, const int32_t* __restrict__ base,
                                                      int32_t* __restrict__ marker) {
    __shared__ float s_obs[OBS_MAX * OBS_STRIDE];
    stage(a, s_obs);
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= a.nz || j >= a.ny) return;
    const size_t c = ((size_t)i * a.ny + j) * a.nz + k;
    marker[c] = last_inside(a, s_obs, i, j, k) >= 0 ? 2 : base[c];
}

// Faces of component c on plane i (grid z = c * (nx + 1) + i).
__global__ __launch_bounds__(GF_WG) void k_obs_faces(ObsArgs a, float* __restrict__ su,
                                                     float* __restrict__ sv,
                                                     float* __restrict__ sw) {
    __shared__ float s_obs[OBS_MAX * OBS_STRIDE];
    stage(a, s_obs);
    const int c = blockIdx.z / (a.nx + 1), i = blockIdx.z % (a.nx + 1);
    const int fx = a.nx + (c == 0), fy = a.ny + (c == 1), fz = a.nz + (c == 2);
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= fx || j >= fy || k >= fz) return;
    // the face's two cells: (i,j,k) and the one below along axis c
    const int b_hi = last_inside(a, s_obs, i, j, k);
    const int b_lo = last_inside(a, s_obs, i - (c == 0), j - (c == 1), k - (c == 2));
    const int b = max(b_hi, b_lo);
    const float vel = b >= 0 ? s_obs[OBS_STRIDE * b + 7 + c] : 0.0f;
    float* f = c == 0 ? su : (c == 1 ? sv : sw);
    f[((size_t)i * fy + j) * fz + k] = vel;
}

}  // namespace

extern "C" int gf_anim_obstacles(const float* obs, int n_obs, const int32_t* base_marker,
                                 int32_t* marker, float* solid_u, float* solid_v, float* solid_w,
                                 int nx, int ny, int nz, double dx, void* stream) {
    GF_REQUIRE(n_obs >= 0 && n_obs <= OBS_MAX, "gf_anim_obstacles: %d obstacles (max %d)", n_obs,
               OBS_MAX);
    GF_REQUIRE(nx > 0 && ny > 0 && nz > 0 && dx > 0.0, "gf_anim_obstacles: grid");
    GF_REQUIRE(nx + 1 <= 21845 && ny + 1 <= 4 * 65535, "gf_anim_obstacles: grid %d x %d", nx, ny);
    GF_REQUIRE(base_marker && marker && solid_u && solid_v && solid_w && (n_obs == 0 || obs),
               "gf_anim_obstacles: null buffer");
    ObsArgs a;
    a.obs = obs; a.n_obs = n_obs; a.nx = nx; a.ny = ny; a.nz = nz; a.dx = dx;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_obs_marker, dim3((nz + 63) / 64, (ny + 3) / 4, nx), dim3(GF_WG), 0, st, a,
                       base_marker, marker);
    hipLaunchKernelGGL(k_obs_faces, dim3((nz + 1 + 63) / 64, (ny + 1 + 3) / 4, 3 * (nx + 1)),
                       dim3(GF_WG), 0, st, a, solid_u, solid_v, solid_w);
    GF_LAUNCH_CHECK("animated obstacles");
    return 0;
}
