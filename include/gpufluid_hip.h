/*
 * gpufluid_hip.h — C ABI of libgpufluid_hip.so, the MI355X (gfx950) FLIP/PIC/APIC
 * substep hot path.
 *
 * Conventions (every entry point):
 *   - returns 0 on success, a negative GF_E* code on a bad argument, or a positive
 *     hipError_t when a launch failed; gf_last_error() then holds a message
 *     (thread-local, valid until the next failing call on the same thread);
 *   - pointers are device pointers to C-order contiguous buffers owned by the
 *     caller, except where a parameter is documented as host memory;
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream);
 *   - no entry point allocates device memory; entry points whose name ends in
 *     `_sync` synchronise `stream` (they read scalars back to the host), every
 *     other entry point only enqueues work on `stream`.
 *
 * Grid layout follows the reference MAC grid (src/gpufluid/solvers/solver3d.py:1-5):
 *   u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1), cell arrays (nx,ny,nz), k fastest;
 *   marker int32: 0 air, 1 fluid, 2 solid.  Particle arrays are (n,3) FP32 (vec3)
 *   and (n,9) FP32 for the APIC affine matrix (row-major mat33).
 *
 * Scene entry points (gf_scene_*) run the fused, tile-based substep on a
 * gf_scene_t descriptor whose buffers the caller allocates with the sizes given
 * by gf_scene_sizes().  Block entry points (gf_k3_*) are one-to-one replacements
 * of the reference's dense Warp kernels and keep their argument meaning.
 */
#ifndef GPUFLUID_HIP_H
#define GPUFLUID_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define GF_OK 0
#define GF_EINVAL (-1)   /* bad argument (null pointer, size, mode) */
#define GF_ERANGE (-2)   /* a size exceeds a documented limit */
#define GF_ESTATE (-3)   /* descriptor inconsistent with the call */

/* ---- library ---------------------------------------------------------- */

/* ABI version (major*10000 + minor*100 + patch). No reference counterpart. */
int gf_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* gf_last_error(void);

/* Select `device` for the calling thread and report its compute-unit count
 * (host int* cus, may be NULL). No reference counterpart
 * (replaces wp.init/device selection, src/gpufluid/primitives/runtime.py:13-24). */
int gf_init(int device, int* cus);

/* Launch-shape tunables, process-wide, read at every call (0 = built-in choice).
 * Keys: "pcg_workgroups" — persistent-grid size of the PCG kernels (default
 * 1024, maximum 2048); "pcg_variant" — 1: two kernels per iteration (standard
 * PCG), 2: one kernel per iteration (Chronopoulos-Gear form), 0 (default):
 * the one-kernel form below 3072 active tiles, else the two-kernel form;
 * "pcg_x_every" — two-kernel form: n = 1..4 (0 or above 4: 4, the default):
 * every n-th update kernel, and a check chunk's last, adds the pending terms
 * alpha p of the iterate x (the same additions in the same order, so the
 * results are bit-identical; it saves n-1 reads and writes of x per n
 * iterations, for n-1 reads of the pending directions).
 * Used by the autotuner (gpufluid_amd/autotune.py).  No reference
 * counterpart (Warp picks its own launch shapes). */
int gf_tune_set(const char* key, int value);
int gf_tune_get(const char* key, int* value);

/* ---- scene descriptor -------------------------------------------------- */

/* Tile edge of the compacted pressure storage and of the P2G/G2P work units. */
#define GF_TILE 8
#define GF_TILE_CELLS 512

/* Marker codes (solver3d.py:1-5). */
#define GF_AIR 0
#define GF_FLUID 1
#define GF_SOLID 2

/* Transfer modes (FlipSolver3D.transfer_mode, solver3d.py:2395). */
#define GF_TRANSFER_FLIP 0   /* FLIP/PIC blend by flip_blend, solver3d.py:1414-1463 */
#define GF_TRANSFER_APIC 1   /* APIC, solver3d.py:1475-1732 */

/* Normalisation rule of gf_scene_attr_transfer. */
#define GF_ATTR_COLOR 0      /* value * (1/w), solver3d.py:607-631 */
#define GF_ATTR_SCALAR 1     /* value / w, solver3d.py:710-727 */

/* n_active argument of gf_scene_jacobi, gf_scene_gsrb and gf_scene_pcg_enqueue
 * (one scene): take the active tile count from s->n_active[0] on the device,
 * as written by gf_scene_sort_sync, instead of a host value.  The launches
 * are then sized for every tile of the grid and do not depend on the count,
 * so a whole substep can be captured in a hipGraph and replayed while the
 * fluid moves (the reference's device n_active, solver3d.py:3369-3385). */
#define GF_NACT_DEVICE (-1)

/* Pressure precision of gf_scene_pcg_sync. */
#define GF_PCG_MIXED 0       /* FP32 vectors, FP64 dots, FP64 residual replacement */
#define GF_PCG_FP64 1        /* plain FP64 PCG */

typedef struct gf_scene {
    /* grid */
    int32_t nx, ny, nz;          /* cells */
    int32_t tnx, tny, tnz;       /* tiles per axis = ceil(n/8) */
    float dx;                    /* cell size as FP32 */
    float dom[3];                /* FP32(n*dx) per axis */
    /* particles (stored in cell-sorted order; pid maps slot -> particle id) */
    int32_t n;                   /* particle count */
    int32_t cap;                 /* allocated particle slots (>= n) */
    float* pos;  float* vel;  float* affine_c;  int32_t* pid;    /* current */
    float* pos_b; float* vel_b; float* affine_c_b; int32_t* pid_b; /* sort target */
    /* sort workspace */
    int32_t* keys;               /* (cap) flat cell key per particle */
    int32_t* tmp;                /* (cap) unstable in-cell slots */
    int32_t* counts;             /* (C) per-cell counts; zero between calls */
    int32_t* cell_start;         /* (C+1) exclusive prefix of counts */
    int32_t* scan_part;          /* (gf_scene_sizes.scan_parts) */
    /* tiles */
    int32_t* tile_occ;           /* (ntiles) bit 0: the tile holds a fluid cell; 1: a particle; 2: a solid cell */
    int32_t* tile_map;           /* (ntiles) compact slot or -1 */
    int32_t* active;             /* (ntiles) tile id per compact slot */
    int32_t* n_active;           /* (4) [0] active tiles, [1] fluid cells */
    int32_t* nbr;                /* (ntiles*6) neighbour slots -x,+x,-y,+y,-z,+z or -1 */
    /* dense MAC grid */
    int32_t* marker;             /* (nx,ny,nz) */
    float* u; float* v; float* w;        /* velocities */
    float* us; float* vs; float* ws;     /* saved pre-pressure velocities */
    const float* solid_u; const float* solid_v; const float* solid_w; /* NULL = 0 */
    /* compact pressure storage, (ntiles*512) entries, slot-major */
    uint8_t* flags;              /* bit3 fluid, bits0-2 non-solid neighbour count */
    float* b;                    /* divergence (reference scaling) */
    float* x32; float* r32; float* p32; float* ap32;  /* FP32 PCG vectors (r32/r64 hold z = r/max(diag,1));
                                                         the two-kernel form keeps the cells' flags at 4 bits
                                                         per cell in ap32 (ap64 in FP64), which it does not
                                                         otherwise use */
    double* x64; double* r64; double* p64; double* ap64; /* FP64 vectors */
    float* pc;  float* ptmp;     /* pressure result / Jacobi ping-pong */
    double* partials;            /* (gf_scene_sizes.partials) per-workgroup partial sums */
    double* scalars;             /* (gf_scene_sizes.scalars) PCG solver log */
    float* p32b; double* p64b;   /* PCG work vector (compact) */
    float* cg32; double* cg64;   /* 3 more PCG work vectors, 3*(ntiles*512) entries each */
    /* batching: several independent scenes packed into this grid, each inside
     * its own solid shell at a tile-aligned offset (n_scenes <= 1: one scene,
     * the fields below unused).  Particles keep their own scene's world
     * units: scene s's particles are the ids [scene_pstart[s],
     * scene_pstart[s+1]), and every per-particle quantity (cell, weights,
     * samples, clamp) is computed with that scene's dx, so each scene's
     * binning and transfers are those of the scene run alone, shifted by its
     * integer cell offset.  Descriptor dx is unused then. */
    int32_t n_scenes;
    const int32_t* tile_scene;   /* (ntiles) scene of each tile, -1 outside every scene */
    const int32_t* tile_order;   /* (ntiles) tile ids in scene-major order (compaction order) */
    const float* scene_par;      /* (n_scenes*GF_SCENE_PAR) per scene, see below */
    int32_t* scene_slot;         /* (n_scenes+1) first compact slot of each scene (sort output) */
    int32_t* pcg_part;           /* (gf_scene_sizes.pcg_part + 2*n_scenes) PCG workgroup partition */
    const int32_t* scene_pstart; /* (n_scenes+1) first particle id of each scene */
} gf_scene_t;

/* Entries of scene_par per scene: [0] g*dt, [1..3] clamp lo x,y,z, [4..6]
 * clamp hi x,y,z (world units: 1.001 dx and n dx - 1.001 dx, in FP32 as the
 * single-scene kernels form them), [7] dx (FP32), [8..10] cell offset of the
 * scene in the packed grid, [11] cells per axis, [12..15] unused. */
#define GF_SCENE_PAR 16

typedef struct gf_scene_size_info {
    int64_t cells;        /* C = nx*ny*nz */
    int64_t ntiles;       /* tnx*tny*tnz */
    int64_t compact;      /* ntiles*512 */
    int64_t scan_parts;   /* entries of scan_part */
    int64_t partials;     /* entries of partials */
    int64_t scalars;      /* entries of scalars (one scene; batched: 48*n_scenes + 64) */
    int64_t pcg_part;     /* entries of pcg_part (one scene; batched: add n_scenes) */
} gf_scene_size_info;

/* Buffer sizes for a grid (host out-pointer). */
int gf_scene_sizes(int nx, int ny, int nz, gf_scene_size_info* out);

/* Bin particles to cells and sort them stably by flat cell key into the *_b
 * arrays; build per-cell ranges, the marker (walls/obstacles kept), the tile
 * map, the compact flags and the tile neighbour table. Writes the active tile
 * count to n_active[0] and the fluid cell count to n_active[1], and returns
 * both in the host array n_active_host[2] (sync; NULL skips the readback).
 * Follows solver3d.py:41-65 (marker clear), 96-101 (cell binning/marking),
 * sim/reseed.py:72-75,114 (flat key, stable in-cell order), 953-969 (8^3 tiles). */
int gf_scene_sort_sync(gf_scene_t* s, int with_affine, int* n_active_host, void* stream);

/* Fused P2G + normalize + gravity + solid BC + divergence over 8^3 tiles
 * (sorted particles, LDS accumulation, no global float atomics).
 * Replaces k3_clear_grid/k3_p2g|k3_p2g_apic/k3_normalize/k3_add_gravity/
 * k3_enforce_solid_bc/k3_compute_divergence (solver3d.py:41-156, 758-799,
 * 846-871, 1475-1588) as launched by _step_impl (solver3d.py:3490-3559).
 * Writes u,v,w (post BC), us,vs,ws (post normalize), b (compact div). */
int gf_scene_p2g(const gf_scene_t* s, int transfer_mode, float gravity, float dt,
                 float rho, void* stream);

/* gf_scene_p2g in two launches: each tile scatters only its own cells'
 * particles (each particle once) into the region they reach and stores it in
 * `stage`; then each tile sums the staged partials of the tiles holding each
 * of its faces (at most two per axis, fixed order) and finalises as
 * gf_scene_p2g.  stage / has: gf_scene_p2g_stage_size floats / int32.
 * Reads s->tile_occ as the last gf_scene_sort_sync left it. */
int gf_scene_p2g_stage_size(int nx, int ny, int nz, int64_t* stage_floats, int64_t* flag_ints);
int gf_scene_p2g_staged(const gf_scene_t* s, int transfer_mode, float gravity, float dt,
                        float rho, float* stage, int32_t* has, void* stream);

/* Raw P2G accumulation into u,v,w (+=) and uw,vw,ww (+=) from the sorted
 * particles: the reference block k3_p2g / k3_p2g_apic (solver3d.py:72-114,
 * 1475-1588) for callers that launch P2G alone. */
int gf_scene_p2g_raw(const gf_scene_t* s, int transfer_mode, float* uw, float* vw,
                     float* ww, void* stream);

/* Jacobi pressure iterations on the compact active tiles, p = 0 initially;
 * result in s->pc. Follows k3_jacobi_pressure (solver3d.py:878-938) and its
 * driver (solver3d.py:3570-3600).  n_active may be GF_NACT_DEVICE. */
int gf_scene_jacobi(const gf_scene_t* s, int n_active, int iters, void* stream);

/* `iters` red-black Gauss-Seidel sweeps (red = (i+j+k) even, then black) of
 * A p = -div in place on the compact tiles from p = 0 (the pressure_solver
 * "gsrb" branch of solver3d.py:3603-3630, sweep k3_gauss_seidel_rb
 * 1204-1267; tile-compact like the reference's per-tile variant
 * k3_gauss_seidel_rb_per_tile).  Result in s->pc.  n_active may be
 * GF_NACT_DEVICE. */
int gf_scene_gsrb(const gf_scene_t* s, int n_active, int iters, void* stream);

/* Diagonal-preconditioned CG for A p = rhs_sign * div on the compact tiles
 * (A = the 7-point operator of k3_apply_A, solver3d.py:1743-1797; Jacobi
 * preconditioner of k3_compute_diag, 1803-1823; driver 2917-3043).
 * precision GF_PCG_MIXED or GF_PCG_FP64; stops when ||r||_2 < tol*||b||_2
 * (mixed: on the FP64 true residual) or after max_iter iterations; the host
 * reads the scalars every `check_every` iterations. rhs_sign = -1 solves the
 * Jacobi-consistent pressure, +1 reproduces the reference PCG's sign.
 * Result in s->pc.  Host outputs (may be NULL): iterations, true relative
 * residual ||b-Ap||/||b|| (FP64; mixed mode only, else the iterated one),
 * number of FP64 residual replacements, and kernel_ms[4]: GPU execution time
 * (HIP events on the launches' dispatch packets on `stream`) of the
 * direction+SpMV (or one-kernel iteration) kernel [0] and the update kernel
 * [1], summed over a sample of one iteration per check_every-chunk (its
 * position in the chunk cycling over the calls of the process), [2] the
 * number of sampled iterations, and [3] the summed time of the residual
 * replacements that ran; NULL disables timing. */
int gf_scene_pcg_sync(const gf_scene_t* s, int n_active, int precision, double tol,
                      int max_iter, int check_every, int rhs_sign, int* iters,
                      double* rel_residual, int* replacements, double* kernel_ms,
                      void* stream);

/* The same solve without any host synchronisation (so it can be captured in
 * a hipGraph): enqueues the setup when first_chunk == 0, then n_chunks chunks
 * of check_every iterations (each with its chunk-end decision), then the
 * result into s->pc.  Scenes that stop early pass through the remaining
 * launches.  A later call with first_chunk = the chunks already enqueued
 * resumes the same solve.  Results: gf_scene_pcg_stats.  For one scene,
 * n_active may be GF_NACT_DEVICE: tiles are split over workgroups as with the
 * host count, and results are bit-identical to it when the grid has fewer than
 * 3072 tiles or the PCG form is pinned (gf_tune_set "pcg_variant"); otherwise
 * the form is chosen from the grid's tile count rather than the active one.
 * The reference's graph path (solver3d.py:3361-3419) runs its PCG kernels
 * under an on-device stop flag in the same way. */
int gf_scene_pcg_enqueue(const gf_scene_t* s, int n_active, int precision, double tol,
                         int max_iter, int check_every, int rhs_sign, int first_chunk,
                         int n_chunks, void* stream);

/* Per-scene results of the last gf_scene_pcg_sync on this descriptor (sync):
 * iterations, true relative residual and FP64 residual replacements, one
 * entry per scene (max(1, n_scenes)); any output may be NULL.  For packed
 * scenes gf_scene_pcg_sync itself reports the maximum iterations and residual
 * and the total replacements.  No reference counterpart. */
int gf_scene_pcg_stats(const gf_scene_t* s, int* iters, double* rel_residual, int* replacements,
                       void* stream);

/* gf_scene_pcg_enqueue with no work buffers from the caller, for callers
 * that cannot allocate (code under hipGraph capture, serving loops, harnesses
 * that forbid hipMalloc).  No reference counterpart.  It allocates nothing:
 * the work vectors, partial sums, live solver log and workgroup partition
 * live in one file-scope device array of GF_PCG_NOALLOC_SCRATCH_BYTES
 * (64.3 MiB) per GPU, held until the process ends.  The descriptor supplies
 * only inputs (flags, nbr, b, the tile counts; scene_slot for packed scenes)
 * and outputs (pc, and the result entries of scalars that
 * gf_scene_pcg_stats / gf_scene_pcg_flags read); its partials, pcg_part,
 * x32/x64 and other PCG vectors are not touched.  Limits (GF_ERANGE beyond
 * them): at most GF_PCG_NOALLOC_MAX_TILES active tiles (with
 * GF_NACT_DEVICE: tiles in the grid) and GF_PCG_NOALLOC_MAX_SCENES scenes.
 * Results are bit-identical to gf_scene_pcg_enqueue with the same arguments.
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents (a resumed solve, first_chunk > 0, continues from them): calls
 * that overlap on one GPU race. */
#define GF_PCG_NOALLOC_MAX_TILES 2048
#define GF_PCG_NOALLOC_MAX_SCENES 64
#define GF_PCG_NOALLOC_SCRATCH_BYTES 67453440
int gf_scene_pcg_enqueue_noalloc(const gf_scene_t* s, int n_active, int precision, double tol,
                                 int max_iter, int check_every, int rhs_sign, int first_chunk,
                                 int n_chunks, void* stream);

/* The reference's PCG stop flag and threshold after the last solve on this
 * descriptor (scene 0), written on the device (enqueue only, capturable):
 * done[0] = 1 if the solve met its tolerance (0 for a zero right-hand side),
 * tol_sq[0] = FP32 tol^2 * ||b||^2.  Replaces the values that
 * k3_zero_int_scalar / k3_set_tol_sq / k3_check_converged leave in
 * FlipSolver3D._pcg_done / _pcg_tol_sq (solver3d.py:1916-1942, 2949-2983).
 * Either output may be NULL, not both. */
int gf_scene_pcg_flags(const gf_scene_t* s, float tol, int32_t* done, float* tol_sq,
                       void* stream);

/* Fused pressure-gradient subtract + solid BC + G2P + advection over 8^3
 * tiles: k3_subtract_pressure_grad (solver3d.py:1364-1407), k3_enforce_solid_bc
 * (758-799), k3_g2p_and_advect (1414-1463) or k3_g2p_apic_advect (1591-1732).
 * `pressure` is a compact pressure vector (s->pc). The pre-projection face
 * values are rebuilt from us,vs,ws exactly as gf_scene_p2g produced them
 * (us [+ gravity*dt on v], then the BC), so this call must follow
 * gf_scene_p2g with the same gravity and dt. Updates u,v,w (faces next to a
 * fluid cell) and the sorted particle arrays in place. */
int gf_scene_g2p(const gf_scene_t* s, const float* pressure, int transfer_mode,
                 float gravity, float dt, float rho, float flip_blend, void* stream);

/* gf_scene_g2p for pipelines that change the velocities between P2G and the
 * pressure solve (viscosity S2.13, solver3d.py:3524-3544): the pre-projection
 * faces are read from pre_u/pre_v/pre_w (distinct from u,v,w) instead of
 * being rebuilt from us,vs,ws; FLIP's "old" velocity is read from us,vs,ws as
 * they stand.  Gravity is not added (it is in pre_*). */
int gf_scene_g2p_pre(const gf_scene_t* s, const float* pressure, const float* pre_u,
                     const float* pre_v, const float* pre_w, int transfer_mode, float dt,
                     float rho, float flip_blend, void* stream);

/* k3_compute_divergence (solver3d.py:846-871) of the dense u,v,w into the
 * compact rhs s->b of the active tiles (0 outside fluid cells), with the
 * arithmetic of the fused P2G; for stages between P2G and the pressure
 * solve. */
int gf_scene_divergence(const gf_scene_t* s, float dt, float rho, void* stream);

/* Surface tension (S2.14, _apply_surface_tension solver3d.py:3230-3300 with
 * k3_indicator_from_marker, k3_box_blur_centred, k3_csf_normal,
 * k3_csf_curvature, k3_csf_apply_u/v/w, k3_csf_sum_* and
 * k3_csf_subtract_bias_*_dev, 203-549) on dense u (nx+1,ny,nz),
 * v (nx,ny+1,nz), w (nx,ny,nz+1), in place: chi = fluid indicator blurred
 * `smoothing_passes` times, n = grad chi/(|grad chi|+1e-8), kappa = -div n,
 * face += coeff*kappa_face*grad chi on faces with no solid side (coeff =
 * sigma*dt/max(rho,1e-8)), then each component's mean change over the faces
 * next to fluid is subtracted there.  The caller applies the solid BC after.
 * Workspace sizes from gf_csf_workspace: work (6*C floats: chi, blur
 * ping-pong, normal, curvature), red (doubles). */
int gf_csf_workspace(int nx, int ny, int nz, int64_t* work_floats, int64_t* red_doubles);
int gf_csf_apply(float* u, float* v, float* w, const int32_t* marker, int nx, int ny, int nz,
                 float dx, float coeff, int smoothing_passes, float* work, double* red,
                 void* stream);

/* ---- Particle-count-changing passes between frames (caller order) ------
 * GPU reseed (S2.11.GPU, sim/reseed.py:161-396): gf_reseed_plan_sync bins
 * the n positions (cell = clamp(int(p/dx))), counts per cell, keeps the
 * first max_per_cell particles of each non-solid cell in index order (those
 * in solid cells are dropped when drop_solid, else kept), and computes each
 * cell's emission count max(min_per_cell - count, 0) over cells that are
 * fluid in the marker or hold particles and are not solid.  counts_host[0]
 * = particles kept, counts_host[1] = particles to emit (sync).
 * gf_reseed_compact moves the kept rows of a (n, width) array to the front
 * in index order; gf_reseed_emit writes the emitted positions, cells in flat
 * order each repeated by its count, at (cell + 0.5 + jitter) * dx with
 * jitter (n_emit, 3) from the caller.  ws: gf_reseed_workspace int32 entries,
 * shared by the three calls. */
int gf_reseed_workspace(int n, int nx, int ny, int nz, int64_t* ws_ints);
int gf_reseed_plan_sync(const float* pos, int n, const int32_t* marker, int nx, int ny, int nz,
                        float dx, int min_per_cell, int max_per_cell, int drop_solid,
                        int32_t* ws, int* counts_host, void* stream);
int gf_reseed_compact(const float* in, int width, int n, int nx, int ny, int nz,
                      const int32_t* ws, float* out, void* stream);
int gf_reseed_emit(int n, int nx, int ny, int nz, float dx, const int32_t* ws,
                   const float* jitter, float* out, void* stream);

/* Outflow compaction (D4.7.GPU, k3_mark_alive_outflows / k3_scatter_alive,
 * solver3d.py:1946-1982, driver 2659-2730): alive[i] = 0 for particles inside
 * any of the n_box boxes (lo x,y,z, hi x,y,z; bounds inclusive), prefix =
 * exclusive scan of alive (n+1 entries), part = ceil(n/4096)+1 scratch;
 * *n_alive_host = survivors (sync).  gf_compact_rows then moves the alive
 * rows of a (n, width) array to prefix[i]. */
int gf_outflow_mark_sync(const float* pos, int n, const float* boxes, int n_box, int32_t* alive,
                         int32_t* prefix, int32_t* part, int* n_alive_host, void* stream);
int gf_compact_rows(const float* in, int width, int n, const int32_t* alive,
                    const int32_t* prefix, float* out, void* stream);

/* Analytic animated obstacles for the per-frame hook (D4.6 in prepare_frame,
 * solver3d.py:3918-3940 with _build_anim_sdf_per_obstacle /
 * _stamp_solid_vel_from_sdf 2537-2596, SDFs primitives/sdf.py:40-66).
 * obs: n_obs (<= 64) rows of 10 floats: kind (0 sphere, 1 box, 2 cylinder
 * along y), centre x,y,z, params (sphere: radius; box: half sizes x,y,z;
 * cylinder: radius, half height), velocity x,y,z.  marker = base_marker with
 * every cell whose SDF at the FP32 cell centre is <= 0 set solid; every face
 * of solid_u/v/w gets the velocity of the last obstacle (list order) that
 * contains one of its two cells, else 0.  dx is the exact cell size. */
int gf_anim_obstacles(const float* obs, int n_obs, const int32_t* base_marker, int32_t* marker,
                      float* solid_u, float* solid_v, float* solid_w, int nx, int ny, int nz,
                      double dx, void* stream);

/* Solid cells inside a closed triangle mesh (D4.3.GPU, schemes/mesh_marker.py:
 * 39-80, 197-236), as prepare_frame applies animated mesh obstacles
 * (solver3d.py:2598-2631).  tris: n_tris triangles of 9 floats (v0, v1, v2).
 * A cell is inside when the ray from its FP32 centre ((i + 0.5) * dx rounded
 * once) along +x crosses an odd number of triangles (the reference's
 * Moller-Trumbore test with direction (1,0,0), t > 1e-6, |det| > 1e-12).
 * gf_mesh_workspace: int32 entries of the workspace for the grid.
 * gf_mesh_bin_sync: bins the triangles into the grid's (j, k) columns and
 * returns the length of the column lists (sync: one 8-byte readback).
 * gf_mesh_mark: with the same triangles and workspace and a list of that
 * length, sets marker = 2 on every inside cell that is not yet 2 and, when
 * the three face arrays are given (else all NULL), writes (vx, vy, vz) into
 * the six faces of each such cell (k3_write_solid_face_vel, solver3d.py:
 * 816-836).  The binning changes no result: a crossing count's parity does
 * not depend on the order of the triangles. */
int gf_mesh_workspace(int nx, int ny, int nz, int64_t* ws_ints);
int gf_mesh_bin_sync(const float* tris, int n_tris, int nx, int ny, int nz, double dx, int32_t* ws,
                     int64_t* n_entries, void* stream);
int gf_mesh_mark(const float* tris, int n_tris, int nx, int ny, int nz, double dx, int32_t* ws,
                 int32_t* list, int64_t n_entries, int32_t* marker, float* solid_u, float* solid_v,
                 float* solid_w, float vx, float vy, float vz, void* stream);

/* Per-particle attribute transfer at the end of a substep (S2.15 colour,
 * solver3d.py:3333-3359 with k3_p2g_color / k3_normalize_color /
 * k3_g2p_color 562-664; S2.18 scalar, 3304-3330 with k3_p2g_scalar /
 * k3_normalize_scalar / k3_g2p_scalar 670-755): trilinear scatter of
 * `channels` (1..3) values per particle onto the cell-centred grid at the
 * current (sorted) positions, normalisation by the deposited weight (mode
 * GF_ATTR_COLOR or GF_ATTR_SCALAR), and trilinear gather back where the
 * gathered weight exceeds 1e-6.  `attr` is (n, channels) FP32 in particle-id
 * order (slot p of the scene holds particle pid[p]) and is updated in place.
 * Workspace: acc (channels+1)*C int64 fixed-point sums, grid (channels+1)*C
 * FP32 (normalised channels, then the weight), scratch 4 uint32; C = the
 * cell count.  Sums are formed in fixed point with integer atomics, so the
 * result is independent of the order of arrival (the reference's FP32
 * atomics are not); exact while each 2x2x2 cell block holds fewer than
 * 32768 particles. */
int gf_scene_attr_transfer(const gf_scene_t* s, float* attr, int channels, int mode,
                           int64_t* acc, float* grid, uint32_t* scratch, void* stream);

/* Scatter a compact cell vector (e.g. s->pc, s->b) into a dense (nx,ny,nz)
 * array; cells of inactive tiles get 0. No reference counterpart. */
int gf_scene_compact_to_dense(const gf_scene_t* s, const float* compact, float* dense,
                              void* stream);

/* Reorder a sorted per-slot particle array (n, width floats) back to particle
 * id order: out[pid[i]] = in[i]. No reference counterpart. */
int gf_scene_unsort(const gf_scene_t* s, const float* in, int width, float* out,
                    void* stream);

/* ---- block entry points: reference kernels, dense arrays -------------- */
/* Each replaces the named Warp kernel with the same argument meaning; the
 * reference's trailing (off_x,off_y,off_z) sub-dense offsets are not taken
 * (dense storage only). nx,ny,nz are cell counts. */

/* k3_clear_grid, solver3d.py:41-65 */
int gf_k3_clear_grid(float* u, float* v, float* w, float* uw, float* vw, float* ww,
                     int32_t* marker, int nx, int ny, int nz, void* stream);
/* marker[cell(p)] = 1 unless solid, solver3d.py:96-101 */
int gf_k3_mark_fluid(const float* pos, int n, int32_t* marker, float dx, int nx, int ny,
                     int nz, void* stream);
/* k3_normalize, solver3d.py:121-142 */
int gf_k3_normalize(float* u, float* v, float* w, const float* uw, const float* vw,
                    const float* ww, float* us, float* vs, float* ws, int nx, int ny,
                    int nz, void* stream);
/* k3_add_gravity, solver3d.py:149-156 (v has (nx,ny+1,nz) entries) */
int gf_k3_add_gravity(float* v, int nx, int ny, int nz, float g, float dt, void* stream);
/* k3_enforce_solid_bc, solver3d.py:758-799 (solid_* may be NULL = 0) */
int gf_k3_enforce_solid_bc(float* u, float* v, float* w, const int32_t* marker,
                           const float* solid_u, const float* solid_v,
                           const float* solid_w, int nx, int ny, int nz, void* stream);
/* k3_compute_divergence, solver3d.py:846-871 */
int gf_k3_compute_divergence(const float* u, const float* v, const float* w, float* div,
                             const int32_t* marker, int nx, int ny, int nz, float dx,
                             float dt, float rho, void* stream);
/* k3_jacobi_pressure, solver3d.py:878-938 */
int gf_k3_jacobi_pressure(const float* p_in, float* p_out, const float* div,
                          const int32_t* marker, int nx, int ny, int nz, void* stream);
/* k3_gauss_seidel_rb, solver3d.py:1204-1267: one in-place half sweep over
 * the cells with (i+j+k) % 2 == color (0 red, 1 black) */
int gf_k3_gauss_seidel_rb(float* p, const float* div, const int32_t* marker, int nx, int ny,
                          int nz, int color, void* stream);
/* k3_jacobi_visc, solver3d.py:178-197: one Jacobi sweep of
 * (I - r Lap) u_out = u_old on one MAC component of shape (sx, sy, sz), reading
 * u_in (distinct from u_out); sx <= 65535, sy <= 262140 */
int gf_k3_jacobi_visc(const float* u_old, const float* u_in, float* u_out, int sx, int sy,
                      int sz, float r, void* stream);
/* k3_jacobi_visc on u (nx+1,ny,nz), v (nx,ny+1,nz) and w (nx,ny,nz+1) in one
 * launch, the substep's S2.13 sweep (solver3d.py:3535-3541); same arithmetic
 * as gf_k3_jacobi_visc.  out_* must not alias in_* or old_*. */
int gf_k3_jacobi_visc3(const float* old_u, const float* old_v, const float* old_w,
                       const float* in_u, const float* in_v, const float* in_w, float* out_u,
                       float* out_v, float* out_w, int nx, int ny, int nz, float r,
                       void* stream);
/* The same sweep on the faces of the box of cells [o, o+b) of the nx*ny*nz
 * grid, with the box as the stencil's array bounds: the reference's
 * k3_jacobi_visc launched on the sub-dense arrays (solver3d.py:178-200,
 * 3530-3541, box from _rebuild_sub_dense 3821-3906).  Faces outside the box
 * are neither read nor written. */
int gf_k3_jacobi_visc3_box(const float* old_u, const float* old_v, const float* old_w,
                           const float* in_u, const float* in_v, const float* in_w,
                           float* out_u, float* out_v, float* out_w, int nx, int ny, int nz,
                           int ox, int oy, int oz, int bx, int by, int bz, float r,
                           void* stream);
/* k3_apply_A, solver3d.py:1743-1797 (done: device int[1], may be NULL) */
int gf_k3_apply_A(const float* x, float* y, const int32_t* marker, const int32_t* done,
                  int nx, int ny, int nz, void* stream);
/* k3_compute_diag, solver3d.py:1803-1823 */
int gf_k3_compute_diag(const int32_t* marker, float* diag, int nx, int ny, int nz,
                       void* stream);
/* k3_apply_invM, solver3d.py:1826-1840 */
int gf_k3_apply_invM(const float* r, const float* diag, float* z, const int32_t* marker,
                     const int32_t* done, int nx, int ny, int nz, void* stream);
/* k3_dot_fluid, solver3d.py:1859-1871: out[0] += sum over fluid a*b (FP64
 * workgroup partials, one FP32 atomic per workgroup) */
int gf_k3_dot_fluid(const float* a, const float* b, float* out, const int32_t* marker,
                    const int32_t* done, int nx, int ny, int nz, void* stream);
/* k3_axpy_devscalar, solver3d.py:1988-2009 */
int gf_k3_axpy_devscalar(const float* a_dev, float sign, const float* x,
                         const float* y_in, float* y_out, const int32_t* marker,
                         const int32_t* done, int nx, int ny, int nz, void* stream);
/* The S2.14 sub-blocks one launch at a time, on dense arrays with the
 * reference's meaning (solver3d.py:203-549): k3_indicator_from_marker,
 * k3_box_blur_centred (src != dst), k3_csf_normal, k3_csf_curvature,
 * k3_csf_apply_u/v/w (comp 0/1/2), k3_csf_sum_u/v/w and
 * k3_csf_subtract_bias_u/v/w_dev.  gf_k3_csf_sum adds the change (after -
 * before) over the component's faces next to fluid, summed in FP64 in a fixed
 * order and rounded to FP32, to sum[0], and the face count to count[0]
 * (device scalars; the reference adds FP32 atomics); `red` is a workspace of
 * gf_k3_csf_sum_workspace doubles.  gf_k3_csf_subtract_bias subtracts
 * sum[0]/count[0] on those faces when count[0] > 0.5. */
int gf_k3_indicator_from_marker(const int32_t* marker, float* chi, int nx, int ny, int nz,
                                void* stream);
int gf_k3_box_blur_centred(const float* src, float* dst, int nx, int ny, int nz, void* stream);
int gf_k3_csf_normal(const float* chi, float* n_x, float* n_y, float* n_z, int nx, int ny, int nz,
                     float dx, void* stream);
int gf_k3_csf_curvature(const float* n_x, const float* n_y, const float* n_z, float* kappa, int nx,
                        int ny, int nz, float dx, void* stream);
int gf_k3_csf_apply(int comp, float* vel, const float* chi, const float* kappa,
                    const int32_t* marker, float coeff, float dx, int nx, int ny, int nz,
                    void* stream);
int gf_k3_csf_sum_workspace(int nx, int ny, int nz, int64_t* red_doubles);
int gf_k3_csf_sum(int comp, const float* before, const float* after, const int32_t* marker,
                  float* sum, float* count, int nx, int ny, int nz, double* red, void* stream);
int gf_k3_csf_subtract_bias(int comp, float* vel, const int32_t* marker, const float* sum,
                            const float* count, int nx, int ny, int nz, void* stream);

/* Colour / scalar transfer one block at a time on the reference's separate
 * FP32 cell grids (S2.15.1-3 k3_p2g_color / k3_normalize_color /
 * k3_g2p_color, solver3d.py:562-664; S2.18.1-3 the scalar forms, 670-755):
 * channels 3 (g0, g1, g2 = r, g, b) or 1 (g0 = value), gw the weight grid;
 * particles in caller order.  gf_k3_p2g_attr adds each cell's exact sum of
 * value*w (and of w), rounded to FP32, to the grids (the reference adds FP32
 * atomics); acc is (channels+1)*nx*ny*nz int64 and scratch 4 words of
 * workspace.  gf_k3_normalize_attr: mode GF_ATTR_COLOR (v * (1/w)) or
 * GF_ATTR_SCALAR (v / w) where w > 1e-8, else 0.  gf_k3_g2p_attr samples the
 * grids at p/dx - 0.5 and updates a particle where the sampled weight
 * exceeds 1e-6. */
int gf_k3_p2g_attr(const float* pos, const float* attr, int channels, int n, float* g0, float* g1,
                   float* g2, float* gw, int nx, int ny, int nz, float dx, int64_t* acc,
                   uint32_t* scratch, void* stream);
int gf_k3_normalize_attr(int mode, int channels, float* g0, float* g1, float* g2, float* gw, int nx,
                         int ny, int nz, void* stream);
int gf_k3_g2p_attr(const float* pos, float* attr, int channels, int n, const float* g0,
                   const float* g1, const float* g2, const float* gw, int nx, int ny, int nz,
                   float dx, void* stream);

/* Active-block bitmask (S2.16, k_mark_active_blocks solver3d.py:952-969):
 * block_active (nbx, nby, nbz) of bs^3 blocks covering the grid; a block with
 * a fluid cell is set to 1, others are left as they are. */
int gf_k3_mark_active_blocks(const int32_t* marker, int32_t* block_active, int nx, int ny, int nz,
                             int bs, int nbx, int nby, int nbz, void* stream);

/* Block-sparse Jacobi sweep (S2.6.4, k3_jacobi_pressure_blocksparse
 * solver3d.py:975-1041): gf_k3_jacobi_pressure with p_out = 0 in blocks whose
 * flag is 0. */
int gf_k3_jacobi_pressure_blocksparse(const float* p_in, float* p_out, const float* div,
                                      const int32_t* marker, const int32_t* block_active, int nx,
                                      int ny, int nz, int bs, int nbx, int nby, int nbz,
                                      void* stream);

/* Per-tile forms (S2.6.4 k3_jacobi_pressure_per_tile solver3d.py:1116-1197,
 * S2.6.5 k3_gauss_seidel_rb_per_tile 1270-1357, S2.6.6 k3_apply_A_per_tile /
 * k3_apply_invM_per_tile / k3_dot_fluid_per_tile / k3_axpy_devscalar_per_tile
 * 2020-2212): the dense kernel of the same name restricted to the cells of
 * the first n_active[0] (a device value) of the max_blocks blocks listed in
 * coords (3 ints per block: block coordinates of bs^3 cells); cells outside
 * those blocks are not touched.  Launches are sized by max_blocks, so the
 * count may change between graph replays. */
int gf_k3_jacobi_pressure_per_tile(const float* p_in, float* p_out, const float* div,
                                   const int32_t* marker, const int32_t* coords, int max_blocks,
                                   int bs, const int32_t* n_active, int nx, int ny, int nz,
                                   void* stream);
int gf_k3_gauss_seidel_rb_per_tile(float* p, const float* div, const int32_t* marker,
                                   const int32_t* coords, int max_blocks, int bs, int color,
                                   const int32_t* n_active, int nx, int ny, int nz, void* stream);
int gf_k3_apply_A_per_tile(const float* x, float* y, const int32_t* marker, const int32_t* coords,
                           int max_blocks, int bs, const int32_t* n_active, const int32_t* done,
                           int nx, int ny, int nz, void* stream);
int gf_k3_apply_invM_per_tile(const float* r, const float* diag, float* z, const int32_t* marker,
                              const int32_t* coords, int max_blocks, int bs,
                              const int32_t* n_active, const int32_t* done, int nx, int ny,
                              int nz, void* stream);
int gf_k3_dot_fluid_per_tile(const float* a, const float* b, float* out, const int32_t* marker,
                             const int32_t* coords, int max_blocks, int bs,
                             const int32_t* n_active, const int32_t* done, int nx, int ny, int nz,
                             void* stream);
int gf_k3_axpy_devscalar_per_tile(const float* a_dev, float sign, const float* x,
                                  const float* y_in, float* y_out, const int32_t* marker,
                                  const int32_t* coords, int max_blocks, int bs,
                                  const int32_t* n_active, const int32_t* done, int nx, int ny,
                                  int nz, void* stream);

/* one-thread scalar kernels, solver3d.py:1874-1939:
 * op 0 k3_zero_scalar(s=a), 1 k3_div_scalar(out=c, numer=a, denom=b),
 * 2 k3_copy_scalar(src=a, dst=c), 3 k3_set_tol_sq(r0=a, tol, out=c),
 * 4 k3_check_converged(r=a, tol_sq=b, done), 5 k3_zero_int_scalar(done) */
int gf_k3_scalar_op(int op, float* a, float* b, float* c, float tol, int32_t* done,
                    void* stream);
/* k3_subtract_pressure_grad, solver3d.py:1364-1407 */
int gf_k3_subtract_pressure_grad(float* u, float* v, float* w, const float* p,
                                 const int32_t* marker, int nx, int ny, int nz, float dx,
                                 float dt, float rho, void* stream);
/* k3_g2p_and_advect, solver3d.py:1414-1463 (dom: host float[3]) */
int gf_k3_g2p_and_advect(float* pos, float* vel, const float* u, const float* v,
                         const float* w, const float* us, const float* vs, const float* ws,
                         int n, float dx, float dt, int nx, int ny, int nz,
                         float flip_blend, const float* dom, void* stream);
/* k3_g2p_apic_advect, solver3d.py:1591-1732 (dom: host float[3]) */
int gf_k3_g2p_apic_advect(float* pos, float* vel, float* affine_c, const float* u,
                          const float* v, const float* w, int n, float dx, float dt,
                          int nx, int ny, int nz, const float* dom, void* stream);
/* k3_vmax_sq_reduce, solver3d.py:2241-2252: out[0] = max(out[0], max |v|^2) */
int gf_k3_vmax_sq_reduce(const float* vel, int n, float* out, void* stream);

/* ---- adjoint of a FLIP/PIC substep (differentiable rollouts) ------------
 * No counterpart in the reference (its solver is not differentiable).  A
 * substep maps particle state (x, v) to (x', v'); these entry points apply its
 * transpose Jacobian, with the fluid/solid masks and the cell binning held
 * fixed.  The descriptor holds the substep's recorded forward state, in the
 * cell-sorted particle order that gf_scene_sort_sync produced:
 *   pos/vel/pid     the substep's inputs after the sort (pid: slot -> particle id)
 *   marker/tile_map the substep's masks and compact tile map
 *   u,v,w           faces after projection and BC (what G2P sampled)
 *   us,vs,ws        normalised faces before gravity (what G2P subtracted)
 * Work arrays (dense face shapes): mu/mv/mw weight sums, au/av/aw and
 * aus/avs/aws adjoint accumulators.  apos/avel (n,3) are the particle
 * adjoints indexed by particle id: the adjoint of (x', v') on entry to
 * gf_adj_g2p, the adjoint of (x, v) after gf_adj_p2g.  Order per substep:
 *   gf_adj_clear, gf_adj_weights, gf_adj_g2p, gf_adj_pressure_rhs,
 *   gf_scene_pcg_sync (same flags/nbr as the forward solve, b = p-bar, same
 *   rhs_sign), gf_adj_faces (b-bar = the PCG result), gf_adj_p2g. */
typedef struct gf_adjoint {
    int32_t nx, ny, nz;
    float dx;
    float dom[3];
    int32_t n;
    const float* pos; const float* vel; const int32_t* pid;
    const int32_t* marker; const int32_t* tile_map;
    const float* u; const float* v; const float* w;
    const float* us; const float* vs; const float* ws;
    float* mu; float* mv; float* mw;
    float* au; float* av; float* aw;
    float* aus; float* avs; float* aws;
    float* apos; float* avel;
} gf_adjoint_t;

/* Zero the nine face work arrays. */
int gf_adj_clear(const gf_adjoint_t* a, void* stream);
/* Rebuild the P2G weight sums m (scatter_face weights, gridmath.py:71-95). */
int gf_adj_weights(const gf_adjoint_t* a, void* stream);
/* Reverse of G2P + advection + clamp (solver3d.py:1414-1463): updates
 * apos/avel in place and scatters into au/av/aw and aus/avs/aws. */
int gf_adj_g2p(const gf_adjoint_t* a, float flip_blend, float dt, void* stream);
/* Reverse of the gradient subtract (solver3d.py:1364-1407): compact p-bar. */
int gf_adj_pressure_rhs(const gf_adjoint_t* a, float dt, float rho, float* pbar, void* stream);
/* Reverse of BC, divergence, gravity and normalize (solver3d.py:121-156,
 * 758-799, 846-871): overwrites au/av/aw with q-bar and aus/avs/aws with m-bar. */
int gf_adj_faces(const gf_adjoint_t* a, const float* bbar, float dt, float rho, void* stream);
/* Reverse of the P2G scatter (solver3d.py:72-114): adds into apos/avel. */
int gf_adj_p2g(const gf_adjoint_t* a, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GPUFLUID_HIP_H */
