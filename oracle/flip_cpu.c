/*
 * flip_cpu.c -- C restatement of the reference FLIP substep on the CPU.
 *
 * TEST INFRASTRUCTURE: this is the CPU baseline that bench.py times beside the
 * GPU run, and a second, independent implementation the tests compare with.
 * It follows the reference arithmetic of src/gpufluid/solvers/solver3d.py
 * (ArkBee/GPUFLUID, gpufluid 0.0.2) on a dense grid, in FP32 like the
 * reference:
 *   clear + marker            solver3d.py:41-65, 96-101
 *   P2G (trilinear scatter)   solver3d.py:72-114, primitives/gridmath.py:71-95
 *   normalize                 solver3d.py:121-142
 *   gravity                   solver3d.py:149-156
 *   solid BC                  solver3d.py:758-799
 *   divergence                solver3d.py:846-871
 *   Jacobi-PCG                solver3d.py:1743-1840, 2917-3043 (x0 = 0, r = div,
 *                             convergence tested every 10 iterations; the result
 *                             is negated to the Jacobi-consistent pressure)
 *   gradient subtract         solver3d.py:1364-1407
 *   G2P + advection (FLIP)    solver3d.py:1414-1463, primitives/gridmath.py:30-65
 *   APIC P2G / G2P            solver3d.py:1475-1588 / 1591-1732 (transfer "apic":
 *                             corners receive v + C row . (x_face - x_p); the
 *                             gather clamps corner indices and rebuilds C)
 * The reference's float atomics are replaced by a race-free schedule: particles
 * are binned into x-slabs of 4 cells and even and odd slabs are scattered in two
 * parallel phases.  Dot products accumulate in double.
 *
 * usage: flip_cpu N DX DT TOL MAX_ITER SUBSTEPS PARTICLES_F32 NP [OUT_F32 [flip|apic]]
 *   PARTICLES_F32 holds NP rows of (x, y, z, vx, vy, vz) float32; OUT_F32 ("-":
 *   none) receives the final positions, velocities and (APIC) the NP x 3 x 3
 *   affine matrices, which start at zero.
 * Prints one JSON line: seconds, Mparticle-substeps/s, PCG iterations.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz;
    float dx, dt, rho, g, flip;
    float dom[3];
    int8_t* marker;
    float *u, *v, *w, *uw, *vw, *ww, *us, *vs, *ws;
    float *p, *div, *r, *z, *d, *ap, *diag;
} Grid;

#define IDX(i, j, k, NY, NZ) (((size_t)(i) * (NY) + (j)) * (NZ) + (k))

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int bin_axis(float p, float dx, int n) {
    float q = p / dx;
    if (!(q > -1.0f)) q = -1.0f; /* NaN and below -> cell 0 */
    if (q > (float)n) q = (float)n;
    return clampi((int)q, 0, n - 1);
}

static void scatter_face(float* f, float* wgt, float val, float fx, float fy, float fz, int NX,
                         int NY, int NZ) {
    int i0 = (int)floorf(fx), j0 = (int)floorf(fy), k0 = (int)floorf(fz);
    float sx = fx - (float)i0, sy = fy - (float)j0, sz = fz - (float)k0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0 || ii >= NX || jj < 0 || jj >= NY || kk < 0 || kk >= NZ) continue;
                float w = (di ? sx : 1.0f - sx) * (dj ? sy : 1.0f - sy) * (dk ? sz : 1.0f - sz);
                size_t o = IDX(ii, jj, kk, NY, NZ);
                f[o] += val * w;
                wgt[o] += w;
            }
}

/* k3_p2g_apic's corner scatter (solver3d.py:1475-1588): the value at a corner
 * is v + C[0]*(xf - px) + C[1]*(yf - py) + C[2]*(zf - pz), with the corner's
 * world position (index + 0 on the component's own axis, + 0.5 on the others)
 * times dx; corners outside the face array are dropped. */
static void scatter_face_apic(float* f, float* wgt, float val, const float* crow, float px,
                              float py, float pz, float dx, int comp, float fx, float fy, float fz,
                              int NX, int NY, int NZ) {
    int i0 = (int)floorf(fx), j0 = (int)floorf(fy), k0 = (int)floorf(fz);
    float sx = fx - (float)i0, sy = fy - (float)j0, sz = fz - (float)k0;
    const float ox = comp == 0 ? 0.0f : 0.5f, oy = comp == 1 ? 0.0f : 0.5f,
                oz = comp == 2 ? 0.0f : 0.5f;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                int ii = i0 + di, jj = j0 + dj, kk = k0 + dk;
                if (ii < 0 || ii >= NX || jj < 0 || jj >= NY || kk < 0 || kk >= NZ) continue;
                float wx = di ? sx : 1.0f - sx, wy = dj ? sy : 1.0f - sy, wz = dk ? sz : 1.0f - sz;
                float wt = wx * wy * wz;
                float fxw = ((float)ii + ox) * dx - px;
                float fyw = ((float)jj + oy) * dx - py;
                float fzw = ((float)kk + oz) * dx - pz;
                float v = val + crow[0] * fxw;
                v = v + crow[1] * fyw;
                v = v + crow[2] * fzw;
                size_t q = IDX(ii, jj, kk, NY, NZ);
                f[q] += v * wt;
                wgt[q] += wt;
            }
}

/* k3_g2p_apic_advect's gather of one component (solver3d.py:1591-1732):
 * corner indices clamped into the face array, velocity = sum w*u, and the
 * affine row C[a] += w*u*(3/dx^2) * (x_face[a] - x_p[a]). */
static float gather_apic(const float* f, float px, float py, float pz, float dx, float inv,
                         int comp, float fx, float fy, float fz, int NX, int NY, int NZ,
                         float* crow) {
    int i0 = (int)floorf(fx), j0 = (int)floorf(fy), k0 = (int)floorf(fz);
    float sx = fx - (float)i0, sy = fy - (float)j0, sz = fz - (float)k0;
    const float off[3] = {comp == 0 ? 0.0f : 0.5f, comp == 1 ? 0.0f : 0.5f,
                          comp == 2 ? 0.0f : 0.5f};
    const float pp[3] = {px, py, pz};
    float vel = 0.0f;
    crow[0] = crow[1] = crow[2] = 0.0f;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                int idx[3] = {clampi(i0 + di, 0, NX - 1), clampi(j0 + dj, 0, NY - 1),
                              clampi(k0 + dk, 0, NZ - 1)};
                float wx = di ? sx : 1.0f - sx, wy = dj ? sy : 1.0f - sy, wz = dk ? sz : 1.0f - sz;
                float wt = wx * wy * wz;
                float val = f[IDX(idx[0], idx[1], idx[2], NY, NZ)];
                vel += wt * val;
                float fu = wt * val * inv;
                for (int a = 0; a < 3; ++a) crow[a] += fu * (((float)idx[a] + off[a]) * dx - pp[a]);
            }
    return vel;
}

static float sample3(const float* f, float fx, float fy, float fz, int nmx, int nmy, int nmz) {
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
    int i1 = i0 + 1 >= nmx ? nmx - 1 : i0 + 1;
    int j1 = j0 + 1 >= nmy ? nmy - 1 : j0 + 1;
    int k1 = k0 + 1 >= nmz ? nmz - 1 : k0 + 1;
    float sx = fminf(fmaxf(fx - (float)i0, 0.0f), 1.0f);
    float sy = fminf(fmaxf(fy - (float)j0, 0.0f), 1.0f);
    float sz = fminf(fmaxf(fz - (float)k0, 0.0f), 1.0f);
#define F(a, b, c) f[IDX(a, b, c, nmy, nmz)]
    float c00 = F(i0, j0, k0) * (1.0f - sx) + F(i1, j0, k0) * sx;
    float c10 = F(i0, j1, k0) * (1.0f - sx) + F(i1, j1, k0) * sx;
    float c01 = F(i0, j0, k1) * (1.0f - sx) + F(i1, j0, k1) * sx;
    float c11 = F(i0, j1, k1) * (1.0f - sx) + F(i1, j1, k1) * sx;
#undef F
    float c0 = c00 * (1.0f - sy) + c10 * sy;
    float c1 = c01 * (1.0f - sy) + c11 * sy;
    return c0 * (1.0f - sz) + c1 * sz;
}

static void* xalloc(size_t n) {
    void* p = calloc(n, 1);
    if (!p) {
        fprintf(stderr, "out of memory (%zu bytes)\n", n);
        exit(3);
    }
    return p;
}

/* A x = diag x - sum over fluid neighbours (solver3d.py:1743-1797). */
static void apply_A(const Grid* G, const float* x, float* y) {
    const int nx = G->nx, ny = G->ny, nz = G->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k, ny, nz);
                if (G->marker[c] != 1) { y[c] = 0.0f; continue; }
                float s = 0.0f, d = 0.0f;
                int m;
                if (i > 0) { m = G->marker[c - (size_t)ny * nz]; if (m != 2) { d += 1; if (m == 1) s += x[c - (size_t)ny * nz]; } }
                if (i < nx - 1) { m = G->marker[c + (size_t)ny * nz]; if (m != 2) { d += 1; if (m == 1) s += x[c + (size_t)ny * nz]; } }
                if (j > 0) { m = G->marker[c - nz]; if (m != 2) { d += 1; if (m == 1) s += x[c - nz]; } }
                if (j < ny - 1) { m = G->marker[c + nz]; if (m != 2) { d += 1; if (m == 1) s += x[c + nz]; } }
                if (k > 0) { m = G->marker[c - 1]; if (m != 2) { d += 1; if (m == 1) s += x[c - 1]; } }
                if (k < nz - 1) { m = G->marker[c + 1]; if (m != 2) { d += 1; if (m == 1) s += x[c + 1]; } }
                y[c] = d * x[c] - s;
            }
}

static double dot_fluid(const Grid* G, const float* a, const float* b, size_t C) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t c = 0; c < C; ++c)
        if (G->marker[c] == 1) s += (double)a[c] * (double)b[c];
    return s;
}

/* Jacobi-preconditioned CG as _pressure_pcg (solver3d.py:2917-3043). */
static int pcg(Grid* G, float tol, int max_iter) {
    const int nx = G->nx, ny = G->ny, nz = G->nz;
    const size_t C = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < C; ++c) {
        G->p[c] = 0.0f;
        G->r[c] = G->div[c];
        float d = 1.0f;
        if (G->marker[c] == 1) {
            int i = (int)(c / ((size_t)ny * nz)), j = (int)((c / nz) % ny), k = (int)(c % nz);
            float s = 0.0f;
            if (i > 0 && G->marker[c - (size_t)ny * nz] != 2) s += 1;
            if (i < nx - 1 && G->marker[c + (size_t)ny * nz] != 2) s += 1;
            if (j > 0 && G->marker[c - nz] != 2) s += 1;
            if (j < ny - 1 && G->marker[c + nz] != 2) s += 1;
            if (k > 0 && G->marker[c - 1] != 2) s += 1;
            if (k < nz - 1 && G->marker[c + 1] != 2) s += 1;
            d = s > 1.0f ? s : 1.0f;
        }
        G->diag[c] = d;
        G->z[c] = G->marker[c] == 1 ? G->r[c] / d : 0.0f;
        G->d[c] = G->z[c];
    }
    float rz_old = (float)dot_fluid(G, G->r, G->z, C);
    float r0 = (float)dot_fluid(G, G->r, G->r, C);
    if (r0 < 1e-30f) return 0;
    const float tol_sq = tol * tol * r0;
    int it;
    for (it = 0; it < max_iter; ++it) {
        apply_A(G, G->d, G->ap);
        float pap = (float)dot_fluid(G, G->d, G->ap, C);
        float alpha = fabsf(pap) < 1e-30f ? 0.0f : rz_old / pap;
#pragma omp parallel for schedule(static)
        for (size_t c = 0; c < C; ++c)
            if (G->marker[c] == 1) {
                G->p[c] = G->p[c] + alpha * G->d[c];
                G->r[c] = G->r[c] + (-alpha) * G->ap[c];
            } else {
                G->p[c] = 0.0f;
                G->r[c] = 0.0f;
            }
        if ((it + 1) % 10 == 0 || it == max_iter - 1) {
            float rr = (float)dot_fluid(G, G->r, G->r, C);
            if (rr < tol_sq) return it + 1;
        }
#pragma omp parallel for schedule(static)
        for (size_t c = 0; c < C; ++c) G->z[c] = G->marker[c] == 1 ? G->r[c] / G->diag[c] : 0.0f;
        float rz_new = (float)dot_fluid(G, G->r, G->z, C);
        float beta = fabsf(rz_old) < 1e-30f ? 0.0f : rz_new / rz_old;
#pragma omp parallel for schedule(static)
        for (size_t c = 0; c < C; ++c)
            G->d[c] = G->marker[c] == 1 ? G->z[c] + beta * G->d[c] : 0.0f;
        rz_old = rz_new;
    }
    return max_iter;
}

static void enforce_bc(Grid* G) {
    const int nx = G->nx, ny = G->ny, nz = G->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k <= nz; ++k) {
                if (j < ny && k < nz) {
                    int l = i == 0 || G->marker[IDX(i - 1, j, k, ny, nz)] == 2;
                    int r = i == nx || G->marker[IDX(i, j, k, ny, nz)] == 2;
                    if (l || r) G->u[IDX(i, j, k, ny, nz)] = 0.0f;
                }
                if (i < nx && k < nz) {
                    int b = j == 0 || G->marker[IDX(i, j - 1, k, ny, nz)] == 2;
                    int t = j == ny || G->marker[IDX(i, j, k, ny, nz)] == 2;
                    if (b || t) G->v[IDX(i, j, k, ny + 1, nz)] = 0.0f;
                }
                if (i < nx && j < ny) {
                    int b = k == 0 || G->marker[IDX(i, j, k - 1, ny, nz)] == 2;
                    int f = k == nz || G->marker[IDX(i, j, k, ny, nz)] == 2;
                    if (b || f) G->w[IDX(i, j, k, ny, nz + 1)] = 0.0f;
                }
            }
}

static int substep(Grid* G, float* pos, float* vel, float* cmat, int np, float tol, int max_iter,
                   int32_t* order, int32_t* slab_start, int nslab, int slab_w) {
    const int nx = G->nx, ny = G->ny, nz = G->nz;
    const size_t C = (size_t)nx * ny * nz;
    const float dx = G->dx, dt = G->dt;
    const size_t nu = (size_t)(nx + 1) * ny * nz, nv = (size_t)nx * (ny + 1) * nz,
                 nw = (size_t)nx * ny * (nz + 1);
    /* clear (41-65) */
    memset(G->u, 0, nu * 4); memset(G->uw, 0, nu * 4);
    memset(G->v, 0, nv * 4); memset(G->vw, 0, nv * 4);
    memset(G->w, 0, nw * 4); memset(G->ww, 0, nw * 4);
    for (size_t c = 0; c < C; ++c)
        if (G->marker[c] != 2) G->marker[c] = 0;
    /* bin particles into x-slabs (counting sort by slab) */
    memset(slab_start, 0, sizeof(int32_t) * (nslab + 1));
    for (int p = 0; p < np; ++p) slab_start[bin_axis(pos[3 * p], dx, nx) / slab_w + 1]++;
    for (int s = 0; s < nslab; ++s) slab_start[s + 1] += slab_start[s];
    {
        int32_t* fill = (int32_t*)xalloc(sizeof(int32_t) * nslab);
        for (int p = 0; p < np; ++p) {
            int s = bin_axis(pos[3 * p], dx, nx) / slab_w;
            order[slab_start[s] + fill[s]++] = p;
        }
        free(fill);
    }
    /* P2G (72-114): even slabs, then odd slabs; a slab's scatter reaches only
     * the neighbouring slabs' boundary faces, never a same-phase slab */
    for (int phase = 0; phase < 2; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int s = phase; s < nslab; s += 2) {
            for (int q = slab_start[s]; q < slab_start[s + 1]; ++q) {
                int p = order[q];
                float x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
                int ci = bin_axis(x, dx, nx), cj = bin_axis(y, dx, ny), ck = bin_axis(z, dx, nz);
                size_t c = IDX(ci, cj, ck, ny, nz);
                if (G->marker[c] != 2) G->marker[c] = 1;
                float fx = x / dx, fy = y / dx, fz = z / dx;
                if (cmat) {
                    const float* cp = cmat + 9 * (size_t)p;
                    scatter_face_apic(G->u, G->uw, vel[3 * p], cp, x, y, z, dx, 0, fx, fy - 0.5f,
                                      fz - 0.5f, nx + 1, ny, nz);
                    scatter_face_apic(G->v, G->vw, vel[3 * p + 1], cp + 3, x, y, z, dx, 1,
                                      fx - 0.5f, fy, fz - 0.5f, nx, ny + 1, nz);
                    scatter_face_apic(G->w, G->ww, vel[3 * p + 2], cp + 6, x, y, z, dx, 2,
                                      fx - 0.5f, fy - 0.5f, fz, nx, ny, nz + 1);
                } else {
                    scatter_face(G->u, G->uw, vel[3 * p], fx, fy - 0.5f, fz - 0.5f, nx + 1, ny, nz);
                    scatter_face(G->v, G->vw, vel[3 * p + 1], fx - 0.5f, fy, fz - 0.5f, nx, ny + 1, nz);
                    scatter_face(G->w, G->ww, vel[3 * p + 2], fx - 0.5f, fy - 0.5f, fz, nx, ny, nz + 1);
                }
            }
        }
    }
    /* normalize (121-142) + gravity (149-156) */
#pragma omp parallel for schedule(static)
    for (size_t f = 0; f < nu; ++f) { float a = G->uw[f] > 1e-8f ? G->u[f] / G->uw[f] : 0.0f; G->u[f] = a; G->us[f] = a; }
#pragma omp parallel for schedule(static)
    for (size_t f = 0; f < nv; ++f) { float a = G->vw[f] > 1e-8f ? G->v[f] / G->vw[f] : 0.0f; G->v[f] = a + G->g * dt; G->vs[f] = a; }
#pragma omp parallel for schedule(static)
    for (size_t f = 0; f < nw; ++f) { float a = G->ww[f] > 1e-8f ? G->w[f] / G->ww[f] : 0.0f; G->w[f] = a; G->ws[f] = a; }
    enforce_bc(G);
    /* divergence (846-871) */
    const float scale = G->rho * dx / dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                size_t c = IDX(i, j, k, ny, nz);
                if (G->marker[c] != 1) { G->div[c] = 0.0f; continue; }
                float d = G->u[IDX(i + 1, j, k, ny, nz)] - G->u[IDX(i, j, k, ny, nz)] +
                          G->v[IDX(i, j + 1, k, ny + 1, nz)] - G->v[IDX(i, j, k, ny + 1, nz)] +
                          G->w[IDX(i, j, k + 1, ny, nz + 1)] - G->w[IDX(i, j, k, ny, nz + 1)];
                G->div[c] = d * scale;
            }
    int iters = pcg(G, tol, max_iter);
    /* reference PCG solves A x = div; the Jacobi-consistent pressure is -x */
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < C; ++c) G->p[c] = -G->p[c];
    /* gradient subtract (1364-1407) */
    const float gs = dt / (G->rho * dx);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
            for (int k = 0; k <= nz; ++k) {
                if (i >= 1 && i <= nx - 1 && j < ny && k < nz) {
                    int ml = G->marker[IDX(i - 1, j, k, ny, nz)], mr = G->marker[IDX(i, j, k, ny, nz)];
                    if ((ml == 1 || mr == 1) && ml != 2 && mr != 2) {
                        float pl = ml == 1 ? G->p[IDX(i - 1, j, k, ny, nz)] : 0.0f;
                        float pr = mr == 1 ? G->p[IDX(i, j, k, ny, nz)] : 0.0f;
                        G->u[IDX(i, j, k, ny, nz)] -= gs * (pr - pl);
                    }
                }
                if (j >= 1 && j <= ny - 1 && i < nx && k < nz) {
                    int mb = G->marker[IDX(i, j - 1, k, ny, nz)], mt = G->marker[IDX(i, j, k, ny, nz)];
                    if ((mb == 1 || mt == 1) && mb != 2 && mt != 2) {
                        float pb = mb == 1 ? G->p[IDX(i, j - 1, k, ny, nz)] : 0.0f;
                        float pt = mt == 1 ? G->p[IDX(i, j, k, ny, nz)] : 0.0f;
                        G->v[IDX(i, j, k, ny + 1, nz)] -= gs * (pt - pb);
                    }
                }
                if (k >= 1 && k <= nz - 1 && i < nx && j < ny) {
                    int mb = G->marker[IDX(i, j, k - 1, ny, nz)], mf = G->marker[IDX(i, j, k, ny, nz)];
                    if ((mb == 1 || mf == 1) && mb != 2 && mf != 2) {
                        float pb = mb == 1 ? G->p[IDX(i, j, k - 1, ny, nz)] : 0.0f;
                        float pf = mf == 1 ? G->p[IDX(i, j, k, ny, nz)] : 0.0f;
                        G->w[IDX(i, j, k, ny, nz + 1)] -= gs * (pf - pb);
                    }
                }
            }
    enforce_bc(G);
    /* G2P + advect (1414-1463; APIC 1591-1732) */
    const float eps = dx * 1.001f;
    if (cmat) {
        const float inv = 3.0f / (dx * dx);
#pragma omp parallel for schedule(static)
        for (int p = 0; p < np; ++p) {
            float px = pos[3 * p], py = pos[3 * p + 1], pz = pos[3 * p + 2];
            float lx = px / dx, ly = py / dx, lz = pz / dx;
            float* cp = cmat + 9 * (size_t)p;
            float nu_ = gather_apic(G->u, px, py, pz, dx, inv, 0, lx, ly - 0.5f, lz - 0.5f,
                                    nx + 1, ny, nz, cp);
            float nv_ = gather_apic(G->v, px, py, pz, dx, inv, 1, lx - 0.5f, ly, lz - 0.5f,
                                    nx, ny + 1, nz, cp + 3);
            float nw_ = gather_apic(G->w, px, py, pz, dx, inv, 2, lx - 0.5f, ly - 0.5f, lz,
                                    nx, ny, nz + 1, cp + 6);
            vel[3 * p] = nu_; vel[3 * p + 1] = nv_; vel[3 * p + 2] = nw_;
            pos[3 * p] = fminf(fmaxf(px + nu_ * dt, eps), G->dom[0] - eps);
            pos[3 * p + 1] = fminf(fmaxf(py + nv_ * dt, eps), G->dom[1] - eps);
            pos[3 * p + 2] = fminf(fmaxf(pz + nw_ * dt, eps), G->dom[2] - eps);
        }
        return iters;
    }
#pragma omp parallel for schedule(static)
    for (int p = 0; p < np; ++p) {
        float lx = pos[3 * p] / dx, ly = pos[3 * p + 1] / dx, lz = pos[3 * p + 2] / dx;
        float pu = sample3(G->u, lx, ly - 0.5f, lz - 0.5f, nx + 1, ny, nz);
        float ou = sample3(G->us, lx, ly - 0.5f, lz - 0.5f, nx + 1, ny, nz);
        float pv = sample3(G->v, lx - 0.5f, ly, lz - 0.5f, nx, ny + 1, nz);
        float ov = sample3(G->vs, lx - 0.5f, ly, lz - 0.5f, nx, ny + 1, nz);
        float pw = sample3(G->w, lx - 0.5f, ly - 0.5f, lz, nx, ny, nz + 1);
        float ow = sample3(G->ws, lx - 0.5f, ly - 0.5f, lz, nx, ny, nz + 1);
        float fb = G->flip;
        float nu_ = fb * (vel[3 * p] + (pu - ou)) + (1.0f - fb) * pu;
        float nv_ = fb * (vel[3 * p + 1] + (pv - ov)) + (1.0f - fb) * pv;
        float nw_ = fb * (vel[3 * p + 2] + (pw - ow)) + (1.0f - fb) * pw;
        vel[3 * p] = nu_; vel[3 * p + 1] = nv_; vel[3 * p + 2] = nw_;
        pos[3 * p] = fminf(fmaxf(pos[3 * p] + nu_ * dt, eps), G->dom[0] - eps);
        pos[3 * p + 1] = fminf(fmaxf(pos[3 * p + 1] + nv_ * dt, eps), G->dom[1] - eps);
        pos[3 * p + 2] = fminf(fmaxf(pos[3 * p + 2] + nw_ * dt, eps), G->dom[2] - eps);
    }
    return iters;
}

int main(int argc, char** argv) {
    if (argc < 9) {
        fprintf(stderr, "usage: %s N DX DT TOL MAX_ITER SUBSTEPS PARTICLES_F32 NP [OUT_F32 [flip|apic]]\n",
                argv[0]);
        return 2;
    }
    const int apic = argc > 10 && strcmp(argv[10], "apic") == 0;
    if (argc > 10 && !apic && strcmp(argv[10], "flip") != 0) {
        fprintf(stderr, "unknown transfer %s\n", argv[10]);
        return 2;
    }
    Grid G;
    memset(&G, 0, sizeof(G));
    const int n = atoi(argv[1]);
    const double dx64 = atof(argv[2]);
    G.nx = G.ny = G.nz = n;
    G.dx = (float)dx64;
    G.dt = (float)atof(argv[3]);
    const float tol = (float)atof(argv[4]);
    const int max_iter = atoi(argv[5]);
    const int substeps = atoi(argv[6]);
    const int np = atoi(argv[8]);
    G.rho = 1.0f;
    G.g = -9.81f;
    G.flip = 0.95f;
    for (int a = 0; a < 3; ++a) G.dom[a] = (float)(n * dx64);
    const size_t C = (size_t)n * n * n, nf = (size_t)(n + 1) * n * n;
    G.marker = (int8_t*)xalloc(C);
    float** faces[] = {&G.u, &G.v, &G.w, &G.uw, &G.vw, &G.ww, &G.us, &G.vs, &G.ws};
    for (int q = 0; q < 9; ++q) *faces[q] = (float*)xalloc(nf * 4);
    float** cells[] = {&G.p, &G.div, &G.r, &G.z, &G.d, &G.ap, &G.diag};
    for (int q = 0; q < 7; ++q) *cells[q] = (float*)xalloc(C * 4);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k)
                if (i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1)
                    G.marker[IDX(i, j, k, n, n)] = 2;
    float* buf = (float*)xalloc((size_t)np * 6 * 4);
    FILE* fp = fopen(argv[7], "rb");
    if (!fp || fread(buf, 4, (size_t)np * 6, fp) != (size_t)np * 6) {
        fprintf(stderr, "cannot read %d particles from %s\n", np, argv[7]);
        return 2;
    }
    fclose(fp);
    float* pos = (float*)xalloc((size_t)np * 12);
    float* vel = (float*)xalloc((size_t)np * 12);
    for (int p = 0; p < np; ++p)
        for (int a = 0; a < 3; ++a) {
            pos[3 * p + a] = buf[6 * p + a];
            vel[3 * p + a] = buf[6 * p + 3 + a];
        }
    free(buf);
    float* cmat = NULL;  /* APIC affine matrices, zero at the start */
    if (apic) {
        cmat = (float*)xalloc((size_t)np * 36);
        memset(cmat, 0, (size_t)np * 36);
    }
    const int slab_w = 4, nslab = (n + slab_w - 1) / slab_w;
    int32_t* order = (int32_t*)xalloc((size_t)np * 4);
    int32_t* slab_start = (int32_t*)xalloc(sizeof(int32_t) * (nslab + 1));
    double t0 = omp_get_wtime();
    int iters = 0;
    for (int s = 0; s < substeps; ++s)
        iters = substep(&G, pos, vel, cmat, np, tol, max_iter, order, slab_start, nslab, slab_w);
    double t1 = omp_get_wtime();
    if (argc > 9 && strcmp(argv[9], "-") != 0) {
        FILE* fo = fopen(argv[9], "wb");
        if (fo) {
            fwrite(pos, 4, (size_t)np * 3, fo);
            fwrite(vel, 4, (size_t)np * 3, fo);
            if (cmat) fwrite(cmat, 4, (size_t)np * 9, fo);
            fclose(fo);
        }
    }
    const double sec = t1 - t0;
    printf("{\"seconds\": %.6f, \"mps\": %.6f, \"iters\": %d, \"threads\": %d, "
           "\"sample\": \"%d substep(s) of %d^3 with %d particles, %s, Jacobi-PCG cap %d\"}\n",
           sec, (double)np * substeps / sec / 1e6, iters, omp_get_max_threads(), substeps, n, np,
           apic ? "APIC" : "FLIP", max_iter);
    return 0;
}
