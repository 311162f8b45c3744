// Cell binning, stable counting sort of particles by flat cell key, and the
// per-substep tile bookkeeping (marker, compact flags, tile neighbour table).
//
// Binning follows solver3d.py:96-99 (clamp(int(p/dx),0,n-1), FP32 divide);
// the flat key i*ny*nz + j*nz + k and the stable in-cell order follow
// sim/reseed.py:72-75 and 114 (np.argsort(kind="stable")).  Only integer
// atomics are used; the in-cell order is fixed afterwards by ranking each
// particle among its cell's members by original index, so the permutation is
// deterministic and equal to a stable sort.
#include "gf_common.h"

namespace {

constexpr int SCAN_ITEMS = 16;
// k_bin / k_slot: each wave takes SORT_SUB consecutive groups of 64
// particles (one per lane each), with all loads of the groups issued before
// the first is used.
constexpr int SORT_SUB = 4;
constexpr int SCAN_CHUNK = GF_WG * SCAN_ITEMS;  // 4096

__global__ void k_fill_i32(int32_t* a, int n, int32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Runs of equal keys among the active lanes of a wave (particles arrive
// nearly cell-sorted, so a cell's members are usually neighbours): the first
// lane of a run gets len = run length (others 0) and every lane gets `lead`,
// the lane index of its run's first lane.  Lanes past the end of the array
// must not call this (the whole wave does at the array tail only when n is a
// multiple of 64; otherwise the guard below keeps exited lanes out).
__device__ __forceinline__ void run_of(int key, int& lead, int& len) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    const bool first = lane == 0 || prev != key;
    const unsigned long long leaders = __ballot(first);
    const unsigned long long active = __ballot(true);
    // highest leader at or below this lane
    const unsigned long long below = leaders & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    lead = 63 - __clzll(below);
    // next leader above this lane (or one past the last active lane)
    const unsigned long long above = leaders & ~((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    const int end = above ? __ffsll(above) - 1 : 64 - __clzll(active);
    len = first ? end - lane : 0;
}

// Packed scenes: a particle is binned in its own scene's grid (its dx and
// size, from the id range it falls in) and the cell shifted by the scene's
// offset, so the cell is the one the scene alone gives it.
__global__ __launch_bounds__(GF_WG) void k_bin(int n, const float* __restrict__ pos,
                                               float dx, int nx, int ny, int nz,
                                               const int32_t* __restrict__ marker,
                                               int32_t* __restrict__ keys,
                                               int32_t* __restrict__ counts,
                                               int32_t* __restrict__ tile_occ, int tny,
                                               int tnz, const int32_t* __restrict__ pid,
                                               const int32_t* __restrict__ scene_pstart,
                                               const float* __restrict__ scene_par,
                                               int n_scenes) {
    const int lane = threadIdx.x & 63;
    const int base = (blockIdx.x * (GF_WG / 64) + (threadIdx.x >> 6)) * (64 * SORT_SUB) + lane;
    float x[SORT_SUB], y[SORT_SUB], z[SORT_SUB];
    int id[SORT_SUB];
#pragma unroll
    for (int j = 0; j < SORT_SUB; ++j) {
        const int p = min(base + 64 * j, n - 1);
        x[j] = pos[3 * p + 0];
        y[j] = pos[3 * p + 1];
        z[j] = pos[3 * p + 2];
        if (scene_pstart) id[j] = pid[p];
    }
#pragma unroll
    for (int j = 0; j < SORT_SUB; ++j) {
        const int p = base + 64 * j;
        if (p >= n) break;  // lanes past the end stay out of run_of
        int ci, cj, ck;
        if (scene_pstart) {
            int lo = 0, hi = n_scenes - 1;  // the last scene whose first id is <= id
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (scene_pstart[mid] <= id[j]) lo = mid;
                else hi = mid - 1;
            }
            const float* sp = scene_par + gf::SCENE_PAR * lo;
            const float sdx = sp[7];
            const int ns = (int)sp[11];
            ci = (int)sp[8] + gf::bin_axis(x[j], sdx, ns);
            cj = (int)sp[9] + gf::bin_axis(y[j], sdx, ns);
            ck = (int)sp[10] + gf::bin_axis(z[j], sdx, ns);
        } else {
            ci = gf::bin_axis(x[j], dx, nx);
            cj = gf::bin_axis(y[j], dx, ny);
            ck = gf::bin_axis(z[j], dx, nz);
        }
        const int key = (ci * ny + cj) * nz + ck;
        keys[p] = key;
        int lead, len;
        run_of(key, lead, len);
        if (len > 0) {
            atomicAdd(&counts[key], len);
            if (marker[key] != 2) tile_occ[((ci >> 3) * tny + (cj >> 3)) * tnz + (ck >> 3)] = 1;
        }
    }
}

__global__ __launch_bounds__(GF_WG) void k_scan_reduce(const int32_t* __restrict__ in, int n,
                                                       int32_t* __restrict__ part) {
    __shared__ int s_wave[GF_WG / GF_WAVE + 1];
    int base = blockIdx.x * SCAN_CHUNK;
    int acc = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int i = base + it * GF_WG + threadIdx.x;
        if (i < n) acc += in[i];
    }
    int total;
    gf::wg_exclusive_scan(acc, s_wave, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Exclusive scan of part[0..nb) in place; part[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_part(int32_t* part, int nb) {
    __shared__ int s_wave[1024 / GF_WAVE + 1];
    int carry = 0;
    for (int base = 0; base < nb; base += blockDim.x) {
        int i = base + threadIdx.x;
        int v = i < nb ? part[i] : 0;
        int total;
        int ex = gf::wg_exclusive_scan(v, s_wave, &total);
        if (i < nb) part[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) part[nb] = carry;
}

__global__ __launch_bounds__(GF_WG) void k_scan_apply(const int32_t* __restrict__ in, int n,
                                                      const int32_t* __restrict__ part,
                                                      int nb, int32_t* __restrict__ out) {
    __shared__ int s_items[SCAN_CHUNK];
    __shared__ int s_wave[GF_WG / GF_WAVE + 1];
    int base = blockIdx.x * SCAN_CHUNK;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int l = it * GF_WG + threadIdx.x;
        int i = base + l;
        s_items[l] = i < n ? in[i] : 0;
    }
    __syncthreads();
    int loc[SCAN_ITEMS];
    int acc = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        loc[it] = acc;
        acc += s_items[threadIdx.x * SCAN_ITEMS + it];
    }
    int total;
    int ex = gf::wg_exclusive_scan(acc, s_wave, &total);
    int off = part[blockIdx.x] + ex;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) s_items[threadIdx.x * SCAN_ITEMS + it] = off + loc[it];
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int l = it * GF_WG + threadIdx.x;
        int i = base + l;
        if (i < n) out[i] = s_items[l];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = part[nb];
}

// A cell whose members all sit in one run of one wave (the common case:
// particles stay in their cell between substeps) needs no slot: k_rank ranks
// them by lane, and the run's leader returns the cell's count to 0.  Other
// cells' members take slots with run-aggregated atomicSub (which also returns
// the count to 0), in no particular order; k_rank orders them.
__global__ __launch_bounds__(GF_WG) void k_slot(int n, const int32_t* __restrict__ keys,
                                                int32_t* __restrict__ counts,
                                                const int32_t* __restrict__ cell_start,
                                                int32_t* __restrict__ tmp) {
    const int lane = threadIdx.x & 63;
    const int base = (blockIdx.x * (GF_WG / 64) + (threadIdx.x >> 6)) * (64 * SORT_SUB) + lane;
    int key[SORT_SUB], s[SORT_SUB], e[SORT_SUB];
#pragma unroll
    for (int j = 0; j < SORT_SUB; ++j) key[j] = keys[min(base + 64 * j, n - 1)];
#pragma unroll
    for (int j = 0; j < SORT_SUB; ++j) {
        s[j] = cell_start[key[j]];
        e[j] = cell_start[key[j] + 1];
    }
#pragma unroll
    for (int j = 0; j < SORT_SUB; ++j) {
        const int p = base + 64 * j;
        if (p >= n) break;  // lanes past the end stay out of run_of
        int lead, len;
        run_of(key[j], lead, len);
        const int run_len = __shfl(len, lead, 64);
        if (run_len == e[j] - s[j]) {
            if (len > 0) counts[key[j]] = 0;
            continue;
        }
        int old = 0;
        if (len > 0) old = atomicSub(&counts[key[j]], len);
        old = __shfl(old, lead, 64);
        tmp[s[j] + old - 1 - (lane - lead)] = p;
    }
}

// Rank each particle among the members of its cell by original index (a stable
// order), then gather its state into the sorted arrays.
__global__ __launch_bounds__(GF_WG) void k_rank(
    int n, const int32_t* __restrict__ keys, const int32_t* __restrict__ cell_start,
    const int32_t* __restrict__ tmp, const float* __restrict__ pos,
    const float* __restrict__ vel, const float* __restrict__ C,
    const int32_t* __restrict__ pid, float* __restrict__ pos_o, float* __restrict__ vel_o,
    float* __restrict__ C_o, int32_t* __restrict__ pid_o) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int key = keys[p];
    int s = cell_start[key], e = cell_start[key + 1];
    // A run of equal keys in this wave that holds every member of the cell
    // (the common case: particles stay in their cell between substeps) is the
    // cell's members in index order already; otherwise count the members with
    // a smaller index.
    int lead, len;
    run_of(key, lead, len);
    const int run_len = __shfl(len, lead, 64);
    int rank = 0;
    if (run_len == e - s)
        rank = (int)(threadIdx.x & 63) - lead;
    else
        for (int q = s; q < e; ++q) rank += (tmp[q] < p);
    int d = s + rank;
    pos_o[3 * d + 0] = pos[3 * p + 0];
    pos_o[3 * d + 1] = pos[3 * p + 1];
    pos_o[3 * d + 2] = pos[3 * p + 2];
    vel_o[3 * d + 0] = vel[3 * p + 0];
    vel_o[3 * d + 1] = vel[3 * p + 1];
    vel_o[3 * d + 2] = vel[3 * p + 2];
    pid_o[d] = pid[p];
    if (C) {
#pragma unroll
        for (int c = 0; c < 9; ++c) C_o[9 * d + c] = C[9 * p + c];
    }
}

// Compact slots of the occupied tiles, in tile order or, for packed scenes,
// in the scene-major order of `order` (so that each scene's slots are
// contiguous; scene_slot[s] = its first slot, scene_slot[n_scenes] = total).
// One workgroup; each thread takes TILE_SCAN_ITEMS consecutive positions of a
// chunk, so a chunk costs one workgroup scan.
constexpr int TILE_SCAN_ITEMS = 16;
__global__ __launch_bounds__(1024) void k_tile_scan(const int32_t* __restrict__ occ,
                                                    int ntiles,
                                                    const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ tile_scene,
                                                    int n_scenes,
                                                    int32_t* __restrict__ scene_slot,
                                                    int32_t* __restrict__ tile_map,
                                                    int32_t* __restrict__ active,
                                                    int32_t* __restrict__ n_active) {
    __shared__ int s_wave[1024 / GF_WAVE + 1];
    constexpr int CHUNK = 1024 * TILE_SCAN_ITEMS;
    int carry = 0;
    for (int base = 0; base < ntiles; base += CHUNK) {
        const int p0 = base + threadIdx.x * TILE_SCAN_ITEMS;
        int tl[TILE_SCAN_ITEMS], o[TILE_SCAN_ITEMS];
        int cnt = 0;
        // every load of the thread's items is issued before the first use
        // (positions past the end read the last tile and count as empty)
#pragma unroll
        for (int i = 0; i < TILE_SCAN_ITEMS; ++i) {
            const int pos = min(p0 + i, ntiles - 1);
            tl[i] = order ? order[pos] : pos;
        }
#pragma unroll
        for (int i = 0; i < TILE_SCAN_ITEMS; ++i) o[i] = occ[tl[i]];
#pragma unroll
        for (int i = 0; i < TILE_SCAN_ITEMS; ++i) {
            o[i] = (p0 + i < ntiles && (o[i] & 1)) ? 1 : 0;
            cnt += o[i];
        }
        int total;
        int ex = carry + gf::wg_exclusive_scan(cnt, s_wave, &total);
#pragma unroll
        for (int i = 0; i < TILE_SCAN_ITEMS; ++i) {
            const int pos = p0 + i;
            if (pos < ntiles) {
                const int t = tl[i];
                tile_map[t] = o[i] ? ex : -1;
                if (o[i]) active[ex] = t;
                if (scene_slot) {
                    const int sc = tile_scene[t];
                    const int prev = pos > 0 ? tile_scene[order[pos - 1]] : -1;
                    if (sc >= 0 && sc != prev) scene_slot[sc] = ex;
                }
                ex += o[i];
            }
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        n_active[0] = carry;
        n_active[1] = 0;  // fluid cells, counted by k_tile_setup
        if (scene_slot) scene_slot[n_scenes] = carry;
    }
}

// Per tile: new marker (solid kept, fluid where a particle sits), compact
// flags (fluid bit + non-solid neighbour count) and the neighbour slot table.
// Every load a thread needs (its two cells' old marker, counts and six
// neighbour markers, and the neighbour tiles' slots) is issued at once.
__global__ __launch_bounds__(GF_WG) void k_tile_setup(
    int nx, int ny, int nz, int tnx, int tny, int tnz, const int32_t* __restrict__ cell_start,
    int32_t* marker, const int32_t* __restrict__ tile_map, uint8_t* __restrict__ flags,
    int32_t* __restrict__ nbr, int32_t* __restrict__ n_active, int32_t* __restrict__ tile_occ) {
    __shared__ int s_wave[GF_WG / GF_WAVE + 1];
    const int ntiles = tnx * tny * tnz;
    const int t = gf::xcd_remap(blockIdx.x, ntiles);
    const int tk = t % tnz, tj = (t / tnz) % tny, ti = t / (tnz * tny);
    const int slot = tile_map[t];
    const int i0 = ti * 8, j0 = tj * 8, k0 = tk * 8;
    const int lj = (threadIdx.x >> 3) & 7, lk = threadIdx.x & 7;
    const int sx = ny * nz;
    int mold[2], cnt[2], nb[2][6];
    bool in[2];
    int cell[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int li = (threadIdx.x >> 6) + 4 * h;
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        in[h] = i < nx && j < ny && k < nz;
        const int c = in[h] ? (i * ny + j) * nz + k : 0;
        cell[h] = c;
        mold[h] = marker[c];
        cnt[h] = cell_start[c + 1] - cell_start[c];
        // neighbours outside the grid read the cell itself and are ignored
        nb[h][0] = marker[i > 0 ? c - sx : c];
        nb[h][1] = marker[i < nx - 1 ? c + sx : c];
        nb[h][2] = marker[j > 0 ? c - nz : c];
        nb[h][3] = marker[j < ny - 1 ? c + nz : c];
        nb[h][4] = marker[k > 0 ? c - 1 : c];
        nb[h][5] = marker[k < nz - 1 ? c + 1 : c];
    }
    int ns = -1;
    if (slot >= 0 && threadIdx.x < 6) {
        const int d = threadIdx.x;
        int a = ti, b = tj, c = tk;
        if (d == 0) a -= 1; else if (d == 1) a += 1;
        else if (d == 2) b -= 1; else if (d == 3) b += 1;
        else if (d == 4) c -= 1; else c += 1;
        if (a >= 0 && a < tnx && b >= 0 && b < tny && c >= 0 && c < tnz)
            ns = tile_map[(a * tny + b) * tnz + c];
    }
    int nfluid = 0, anyp = 0, anys = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int li = (threadIdx.x >> 6) + 4 * h;
        const int i = i0 + li, j = j0 + lj, k = k0 + lk;
        const int l = (li * 8 + lj) * 8 + lk;
        uint8_t f = 0;
        if (in[h]) {
            anyp |= cnt[h] > 0;
            anys |= mold[h] == 2;
            const int m = (mold[h] == 2) ? 2 : (cnt[h] > 0 ? 1 : 0);
            if (m != mold[h]) marker[cell[h]] = m;
            if (m == 1) {
                int d = 0;
                if (i > 0 && nb[h][0] != 2) ++d;
                if (i < nx - 1 && nb[h][1] != 2) ++d;
                if (j > 0 && nb[h][2] != 2) ++d;
                if (j < ny - 1 && nb[h][3] != 2) ++d;
                if (k > 0 && nb[h][4] != 2) ++d;
                if (k < nz - 1 && nb[h][5] != 2) ++d;
                f = (uint8_t)(8 | d);
                ++nfluid;
            }
        }
        if (slot >= 0) flags[(size_t)slot * GF_TILE_CELLS + l] = f;
    }
    int total;
    gf::wg_exclusive_scan(nfluid, s_wave, &total);
    if (threadIdx.x == 0 && total > 0) atomicAdd(&n_active[1], total);
    // bit 1: the tile holds a particle (in any cell, solid ones included);
    // the staged P2G scatters only such tiles.  bit 2: it holds a solid cell.
    const int bits = (__syncthreads_or(anyp) ? 2 : 0) | (__syncthreads_or(anys) ? 4 : 0);
    if (bits && threadIdx.x == 0) tile_occ[t] |= bits;
    if (slot >= 0 && threadIdx.x < 6) nbr[(size_t)slot * 6 + threadIdx.x] = ns;
}

int scan_exclusive(const int32_t* in, int n, int32_t* out, int32_t* part, hipStream_t st) {
    int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(GF_WG), 0, st, in, n, part);
    hipLaunchKernelGGL(k_scan_part, dim3(1), dim3(1024), 0, st, part, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(GF_WG), 0, st, in, n, part, nb, out);
    return gf::check_launch("scan");
}

}  // namespace

namespace gf {
int scan_exclusive_i32(const int32_t* in, int n, int32_t* out, int32_t* part, hipStream_t st) {
    return scan_exclusive(in, n, out, part, st);
}

int scene_check(const gf_scene_t* s, const char* who) {
    if (!s) return set_error(GF_EINVAL, "%s: null scene", who);
    if (s->nx < 1 || s->ny < 1 || s->nz < 1)
        return set_error(GF_EINVAL, "%s: grid %dx%dx%d", who, s->nx, s->ny, s->nz);
    if (s->tnx != (s->nx + 7) / 8 || s->tny != (s->ny + 7) / 8 || s->tnz != (s->nz + 7) / 8)
        return set_error(GF_ESTATE, "%s: tile counts do not match the grid", who);
    if (s->n < 0 || s->n > s->cap)
        return set_error(GF_ESTATE, "%s: n=%d outside [0, cap=%d]", who, s->n, s->cap);
    if (!(s->dx > 0.0f)) return set_error(GF_EINVAL, "%s: dx must be > 0", who);
    if (s->n_scenes < 0) return set_error(GF_EINVAL, "%s: n_scenes %d", who, s->n_scenes);
    if (s->n_scenes > 1 && (!s->tile_scene || !s->tile_order || !s->scene_par ||
                            !s->scene_slot || !s->pcg_part || !s->scene_pstart))
        return set_error(GF_EINVAL, "%s: packed scenes need tile_scene, tile_order, scene_par, "
                                    "scene_slot, pcg_part and scene_pstart", who);
    return 0;
}
}  // namespace gf

extern "C" int gf_scene_sort_sync(gf_scene_t* s, int with_affine, int* n_active_host,
                                  void* stream) {
    int rc = gf::scene_check(s, "gf_scene_sort_sync");
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    GF_REQUIRE(s->keys && s->tmp && s->counts && s->cell_start && s->scan_part,
               "gf_scene_sort_sync: sort workspace not set");
    GF_REQUIRE(s->tile_occ && s->tile_map && s->active && s->n_active && s->nbr && s->flags,
               "gf_scene_sort_sync: tile workspace not set");
    GF_REQUIRE(s->marker, "gf_scene_sort_sync: marker not set");
    if (s->n > 0)
        GF_REQUIRE(s->pos && s->vel && s->pid && s->pos_b && s->vel_b && s->pid_b,
                   "gf_scene_sort_sync: particle buffers not set");
    if (with_affine && s->n > 0)
        GF_REQUIRE(s->affine_c && s->affine_c_b, "gf_scene_sort_sync: affine buffers not set");
    const int C = s->nx * s->ny * s->nz;
    const int ntiles = s->tnx * s->tny * s->tnz;
    hipLaunchKernelGGL(k_fill_i32, dim3((ntiles + GF_WG - 1) / GF_WG), dim3(GF_WG), 0, st,
                       s->tile_occ, ntiles, 0);
    const int npb = (s->n + GF_WG - 1) / GF_WG;
    const int nsb = (s->n + GF_WG * SORT_SUB - 1) / (GF_WG * SORT_SUB);
    if (s->n > 0)
        hipLaunchKernelGGL(k_bin, dim3(nsb), dim3(GF_WG), 0, st, s->n, s->pos, s->dx, s->nx,
                           s->ny, s->nz, s->marker, s->keys, s->counts, s->tile_occ, s->tny,
                           s->tnz, s->pid, s->n_scenes > 1 ? s->scene_pstart : nullptr,
                           s->scene_par, s->n_scenes);
    GF_LAUNCH_CHECK("k_bin");
    rc = scan_exclusive(s->counts, C, s->cell_start, s->scan_part, st);
    if (rc) return rc;
    if (s->n > 0) {
        hipLaunchKernelGGL(k_slot, dim3(nsb), dim3(GF_WG), 0, st, s->n, s->keys, s->counts,
                           s->cell_start, s->tmp);
        hipLaunchKernelGGL(k_rank, dim3(npb), dim3(GF_WG), 0, st, s->n, s->keys,
                           s->cell_start, s->tmp, s->pos, s->vel,
                           with_affine ? s->affine_c : nullptr, s->pid, s->pos_b, s->vel_b,
                           with_affine ? s->affine_c_b : nullptr, s->pid_b);
        GF_LAUNCH_CHECK("k_slot/k_rank");
    }
    const bool packed = s->n_scenes > 1;
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, st, s->tile_occ, ntiles,
                       packed ? s->tile_order : nullptr, packed ? s->tile_scene : nullptr,
                       s->n_scenes, packed ? s->scene_slot : nullptr, s->tile_map, s->active,
                       s->n_active);
    hipLaunchKernelGGL(k_tile_setup, dim3(ntiles), dim3(GF_WG), 0, st, s->nx, s->ny, s->nz,
                       s->tnx, s->tny, s->tnz, s->cell_start, s->marker, s->tile_map,
                       s->flags, s->nbr, s->n_active, s->tile_occ);
    GF_LAUNCH_CHECK("k_tile_scan/k_tile_setup");
    if (n_active_host) {
        int32_t na[2] = {0, 0};
        hipError_t e = hipMemcpyAsync(na, s->n_active, 2 * sizeof(int32_t),
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            return gf::set_error((int)e, "gf_scene_sort_sync: readback: %s",
                                 hipGetErrorString(e));
        n_active_host[0] = na[0];
        n_active_host[1] = na[1];
    }
    return 0;
}
