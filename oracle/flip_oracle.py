"""CPU restatement of the reference FLIP/PIC/APIC substep — TEST INFRASTRUCTURE.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  It restates, in numpy (FP64 by default, or FP32-faithful), the arithmetic
of the reference's Warp kernels in ``src/gpufluid/solvers/solver3d.py`` and
``src/gpufluid/primitives/gridmath.py`` (ArkBee/GPUFLUID, gpufluid 0.0.2);
each function cites the lines it follows.

How each part is pinned (tests/test_oracle_*.py):
  * cell binning / flat key / stable order: closed form clamp(trunc(fp32(p)/
    fp32(dx)),0,n-1) on boundary vectors, brute-force per-particle loop, and
    ``np.argsort(kind="stable")`` (reference sim/reseed.py:75,114);
  * marker: set algebra walls U solid U {cells with particles};
  * P2G + normalize: pure-Python per-particle loop on tiny grids, partition of
    unity of the trilinear weights, uniform-velocity invariance;
  * gravity / BC / divergence / gradient: closed forms on hand-built fields;
  * Jacobi: fixed point equals the scipy.sparse solve of the assembled
    7-point system; 1-cell and 1-D tridiagonal textbook cases;
  * PCG (reference restatement): converges to the spsolve solution with the
    reference's sign (x = -p); iteration counts on tiny cases;
  * sample3 / G2P: trilinear interpolation reproduces affine fields exactly;
    PIC == FLIP with flip_blend 0; clamp box;
  * APIC: affine velocity field is transferred exactly (P2G->G2P round trip).
  * viscosity sweep: constant fields fixed, r = 0 identity, fixed point =
    numpy.linalg.solve of the assembled (I - r L) system, per-entry loop;
  * surface tension: blur = brute-force 27-point mean; normal and curvature
    of a blurred sphere point inward/≈2/R; force balance leaves zero mean
    change; per-face loop of the impulse;
  * attribute transfer: hand-derived closed form on a two-particle line,
    identity for one particle at each cell centre, constants preserved,
    weight sum = particle count, pure-Python per-corner loop.
Functions without such a check say "parity unpinned" in their docstring.
"""
from __future__ import annotations

import numpy as np

try:  # scipy is only needed for the assembled-matrix checks
    import scipy.sparse as _sp
    import scipy.sparse.linalg as _spla
except Exception:  # pragma: no cover
    _sp = None
    _spla = None

AIR, FLUID, SOLID = 0, 1, 2


# ----------------------------------------------------------------- grid set-up
def wall_marker(nx: int, ny: int, nz: int) -> np.ndarray:
    """One-cell solid shell (solver3d.py:2475-2482)."""
    m = np.zeros((nx, ny, nz), dtype=np.int32)
    m[0, :, :] = SOLID
    m[-1, :, :] = SOLID
    m[:, 0, :] = SOLID
    m[:, -1, :] = SOLID
    m[:, :, 0] = SOLID
    m[:, :, -1] = SOLID
    return m


def domain(n: int, dx: float) -> np.float32:
    """Domain extent as the kernels see it: FP32 of the FP64 product n*dx
    (solver3d.py:2382 builds wp.vec3(n*dx, ...))."""
    return np.float32(n * dx)


# --------------------------------------------------------------- cell binning
def cell_index(pos: np.ndarray, dx: float, shape) -> tuple:
    """clamp(int(p/dx), 0, n-1) per axis with an FP32 divide and truncation
    toward zero (solver3d.py:96-99, 1502-1507).  NaN maps to cell 0."""
    p = np.asarray(pos, dtype=np.float32)
    q = p / np.float32(dx)
    out = []
    for ax in range(3):
        n = shape[ax]
        qa = np.nan_to_num(q[:, ax], nan=0.0, posinf=float(n), neginf=-1.0)
        qa = np.clip(qa, -1.0, float(n))
        c = np.trunc(qa).astype(np.int64)
        out.append(np.clip(c, 0, n - 1))
    return tuple(out)


def flat_key(ci, cj, ck, shape) -> np.ndarray:
    """Flat C-order cell key i*ny*nz + j*nz + k (sim/reseed.py:72-75)."""
    _, ny, nz = shape
    return (ci * ny + cj) * nz + ck


def stable_order(key: np.ndarray) -> np.ndarray:
    """Stable counting-sort permutation by key (sim/reseed.py:114)."""
    return np.argsort(key, kind="stable")


def update_marker(marker: np.ndarray, pos: np.ndarray, dx: float) -> np.ndarray:
    """Clear non-solid cells, then mark cells holding a particle fluid unless
    solid (solver3d.py:62-64 then 100-101)."""
    m = marker.copy()
    m[m != SOLID] = AIR
    if len(pos):
        ci, cj, ck = cell_index(pos, dx, m.shape)
        hit = np.zeros(m.shape, dtype=bool)
        hit[ci, cj, ck] = True
        m[hit & (m != SOLID)] = FLUID
    return m


# ------------------------------------------------------------------------ P2G
def _face_shapes(shape):
    nx, ny, nz = shape
    return (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)


def _corners(c):
    """floor and fractional parts of face-space coordinates (gridmath.py:79-80)."""
    i0 = np.floor(c).astype(np.int64)
    return i0, c - i0


def p2g(pos, vel, marker, dx, dtype=np.float64, affine=None):
    """Raw P2G scatter: returns (u, v, w, uw, vw, ww).  FLIP/PIC follows
    k3_p2g (solver3d.py:72-114) with scatter_face (gridmath.py:71-95):
    corners outside the face array are dropped.  With ``affine`` (n,3,3) it
    follows k3_p2g_apic (solver3d.py:1475-1588): each corner receives
    v + C[row]·(x_face - x_p) with world-space face positions."""
    shapes = _face_shapes(marker.shape)
    fields = [np.zeros(s, dtype=dtype) for s in shapes]
    weights = [np.zeros(s, dtype=dtype) for s in shapes]
    if len(pos) == 0:
        return (*fields, *weights)
    if dtype == np.float32:
        p = np.asarray(pos, dtype=np.float32)
        q = p / np.float32(dx)
        half = np.float32(0.5)
        dxc = np.float32(dx)
    else:
        p = np.asarray(pos, dtype=np.float32).astype(np.float64)
        dxc = float(np.float32(dx))
        q = p / dxc
        half = 0.5
    vv = np.asarray(vel).astype(dtype)
    C = None if affine is None else np.asarray(affine).reshape(-1, 3, 3).astype(dtype)
    for comp in range(3):
        c = q.copy()
        for ax in range(3):
            if ax != comp:
                c[:, ax] = q[:, ax] - half
        i0, frac = _corners(c)
        shp = shapes[comp]
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    ii, jj, kk = i0[:, 0] + di, i0[:, 1] + dj, i0[:, 2] + dk
                    ok = ((ii >= 0) & (ii < shp[0]) & (jj >= 0) & (jj < shp[1])
                          & (kk >= 0) & (kk < shp[2]))
                    wx = frac[:, 0] if di else 1 - frac[:, 0]
                    wy = frac[:, 1] if dj else 1 - frac[:, 1]
                    wz = frac[:, 2] if dk else 1 - frac[:, 2]
                    wt = (wx * wy * wz).astype(dtype)
                    val = vv[:, comp]
                    if C is not None:
                        fpos = []
                        for ax, idx in enumerate((ii, jj, kk)):
                            off = 0.0 if ax == comp else 0.5
                            fpos.append((idx.astype(dtype) + dtype(off)) * dxc - p[:, ax])
                        val = (val + C[:, comp, 0] * fpos[0] + C[:, comp, 1] * fpos[1]
                               + C[:, comp, 2] * fpos[2])
                    np.add.at(fields[comp], (ii[ok], jj[ok], kk[ok]), (val * wt)[ok])
                    np.add.at(weights[comp], (ii[ok], jj[ok], kk[ok]), wt[ok])
    return (*fields, *weights)


def p2g_loop(pos, vel, marker, dx):
    """Pure-Python per-particle restatement of scatter_face
    (gridmath.py:71-95) in FP64, for tiny cases; used to pin ``p2g``."""
    shapes = _face_shapes(marker.shape)
    fields = [np.zeros(s) for s in shapes]
    weights = [np.zeros(s) for s in shapes]
    dxc = float(np.float32(dx))
    for n in range(len(pos)):
        q = [float(np.float32(pos[n][a])) / dxc for a in range(3)]
        for comp in range(3):
            c = [q[a] - (0.0 if a == comp else 0.5) for a in range(3)]
            i0 = [int(np.floor(x)) for x in c]
            s = [c[a] - i0[a] for a in range(3)]
            for di in (0, 1):
                for dj in (0, 1):
                    for dk in (0, 1):
                        ii, jj, kk = i0[0] + di, i0[1] + dj, i0[2] + dk
                        shp = shapes[comp]
                        if 0 <= ii < shp[0] and 0 <= jj < shp[1] and 0 <= kk < shp[2]:
                            w = ((s[0] if di else 1 - s[0]) * (s[1] if dj else 1 - s[1])
                                 * (s[2] if dk else 1 - s[2]))
                            fields[comp][ii, jj, kk] += float(vel[n][comp]) * w
                            weights[comp][ii, jj, kk] += w
    return (*fields, *weights)


def normalize(f, wgt):
    """f/w where w > 1e-8 else 0 (solver3d.py:121-142); returns (value, saved copy)."""
    out = np.where(wgt > 1e-8, f / np.where(wgt > 1e-8, wgt, 1), 0).astype(f.dtype)
    return out, out.copy()


def add_gravity(v, g, dt):
    """v += g*dt on every v face (solver3d.py:149-156); g*dt in the field's precision."""
    return (v + v.dtype.type(g) * v.dtype.type(dt)).astype(v.dtype)


def enforce_solid_bc(u, v, w, marker, solid_u=None, solid_v=None, solid_w=None):
    """Faces on the domain edge or next to a solid cell take the obstacle
    velocity (zero when none is given) (solver3d.py:758-799)."""
    u, v, w = u.copy(), v.copy(), w.copy()
    solid = marker == SOLID
    for comp, f, sf in ((0, u, solid_u), (1, v, solid_v), (2, w, solid_w)):
        n = marker.shape[comp]
        mask = np.zeros(f.shape, dtype=bool)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[comp] = 0
        hi[comp] = n
        mask[tuple(lo)] = True
        mask[tuple(hi)] = True
        left = [slice(None)] * 3
        left[comp] = slice(1, n + 1)
        mask[tuple(left)] |= solid               # cell below the face is solid
        right = [slice(None)] * 3
        right[comp] = slice(0, n)
        mask[tuple(right)] |= solid              # cell above the face is solid
        f[mask] = 0 if sf is None else sf[mask]
    return u, v, w


def divergence(u, v, w, marker, dx, dt, rho):
    """(Δu+Δv+Δw)·ρ·dx/dt on fluid cells, else 0 (solver3d.py:846-871);
    summed in the reference's left-to-right order."""
    t = u.dtype.type
    d = (u[1:] - u[:-1]) + v[:, 1:] - v[:, :-1] + w[:, :, 1:] - w[:, :, :-1]
    scale = t(rho) * t(dx) / t(dt)
    return np.where(marker == FLUID, d * scale, 0).astype(u.dtype)


# ------------------------------------------------------------------ pressure
_DIRS = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))


def _shift(a, ax, d, fill):
    """b[c] = a[c + d along ax]; out-of-grid entries = fill."""
    out = np.full_like(a, fill)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if d < 0:
        src[ax] = slice(0, -1)
        dst[ax] = slice(1, None)
    else:
        src[ax] = slice(1, None)
        dst[ax] = slice(0, -1)
    out[tuple(dst)] = a[tuple(src)]
    return out


def neighbour_count(marker):
    """Number of in-grid non-solid neighbours (the diagonal of A,
    solver3d.py:1767-1796, 1815-1821)."""
    d = np.zeros(marker.shape, dtype=np.int32)
    for ax, s in _DIRS:
        nb = _shift(marker, ax, s, SOLID)
        d += (nb != SOLID)
    return d


def _sum_fluid_nb(x, marker):
    s = np.zeros_like(x)
    for ax, d in _DIRS:                       # -x,+x,-y,+y,-z,+z as in the reference
        nbm = _shift(marker, ax, d, SOLID)
        nbx = _shift(x, ax, d, 0)
        s = s + np.where(nbm == FLUID, nbx, 0).astype(x.dtype)
    return s


def jacobi_sweep(p_in, div, marker):
    """One Jacobi sweep p_out = (Σ_fluid-nb p_in − div)/diag on fluid cells,
    0 elsewhere or when diag < 0.5 (solver3d.py:878-938)."""
    diag = neighbour_count(marker).astype(p_in.dtype)
    s = _sum_fluid_nb(p_in, marker)
    ok = (marker == FLUID) & (diag >= 0.5)
    return np.where(ok, (s - div) / np.where(ok, diag, 1), 0).astype(p_in.dtype)


def jacobi(div, marker, iters, dtype=np.float32):
    """Driver of solver3d.py:3570-3600: p = 0, then `iters` sweeps."""
    p = np.zeros(marker.shape, dtype=dtype)
    d = div.astype(dtype)
    for _ in range(iters):
        p = jacobi_sweep(p, d, marker)
    return p


def gsrb_half_sweep(p, div, marker, color):
    """One in-place red-black Gauss-Seidel half sweep (solver3d.py:1204-1267):
    cells with (i+j+k) % 2 == color get (Σ_fluid-nb p − div)/diag on fluid
    cells with diag >= 0.5, else 0; other cells keep their value.  A cell's
    neighbours all have the other colour, so updating the whole colour at
    once equals the sequential in-place loop."""
    ii, jj, kk = np.indices(marker.shape)
    mine = ((ii + jj + kk) % 2) == color
    new = jacobi_sweep(p, div.astype(p.dtype), marker)
    return np.where(mine, new, p).astype(p.dtype)


def gsrb(div, marker, iters, dtype=np.float32):
    """Driver of the "gsrb" pressure branch (solver3d.py:3603-3630): p = 0,
    then `iters` red + black half sweeps."""
    p = np.zeros(marker.shape, dtype=dtype)
    for _ in range(iters):
        p = gsrb_half_sweep(p, div, marker, 0)
        p = gsrb_half_sweep(p, div, marker, 1)
    return p


def apply_A(x, marker):
    """A x = diag·x − Σ_fluid-nb x on fluid cells, 0 elsewhere (solver3d.py:1743-1797)."""
    diag = neighbour_count(marker).astype(x.dtype)
    return np.where(marker == FLUID, diag * x - _sum_fluid_nb(x, marker), 0).astype(x.dtype)


def precond_diag(marker):
    """max(diag, 1) on fluid cells, 1 elsewhere (solver3d.py:1803-1823)."""
    d = neighbour_count(marker).astype(np.float64)
    return np.where(marker == FLUID, np.maximum(d, 1.0), 1.0)


def poisson_matrix(marker):
    """Assembled sparse 7-point matrix over the fluid cells (row order = C
    order of the fluid cells) and the index map; built from index arithmetic,
    independently of the stencil code above."""
    if _sp is None:
        raise RuntimeError("scipy is required for poisson_matrix")
    fluid = np.argwhere(marker == FLUID)
    index = -np.ones(marker.shape, dtype=np.int64)
    index[tuple(fluid.T)] = np.arange(len(fluid))
    rows, cols, vals = [], [], []
    shape = np.array(marker.shape)
    diag = np.zeros(len(fluid))
    for off in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        nb = fluid + np.array(off)
        inside = np.all((nb >= 0) & (nb < shape), axis=1)
        nbc = np.where(inside[:, None], nb, 0)
        m = marker[tuple(nbc.T)]
        nonsolid = inside & (m != SOLID)
        diag += nonsolid
        isf = inside & (m == FLUID)
        rows.append(np.nonzero(isf)[0])
        cols.append(index[tuple(nbc[isf].T)])
        vals.append(-np.ones(isf.sum()))
    rows.append(np.arange(len(fluid)))
    cols.append(np.arange(len(fluid)))
    vals.append(diag)
    A = _sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(len(fluid), len(fluid)))
    return A, fluid, index


def solve_exact(div, marker):
    """FP64 direct solve of A p = −div over the fluid cells (the Jacobi fixed
    point, physically signed); p = 0 off fluid."""
    A, fluid, _ = poisson_matrix(marker)
    p = np.zeros(marker.shape)
    if len(fluid):
        rhs = -div[tuple(fluid.T)].astype(np.float64)
        p[tuple(fluid.T)] = _spla.spsolve(A.tocsc(), rhs)
    return p


def pcg_reference(div, marker, max_iter, tol=1e-4, check_every=10, dtype=np.float32):
    """Literal restatement of FlipSolver3D._pressure_pcg (solver3d.py:2917-3043):
    x0 = 0, r = b = +div (so x = −p, the reference's sign), Jacobi
    preconditioner, α = 0 when |pAp| < 1e-30, β = 0 when |rz| < 1e-30,
    stop when r·r < tol²·r0·r0 evaluated every `check_every` iterations and at
    the last one.  Returns (x, iterations)."""
    t = np.dtype(dtype).type
    fl = marker == FLUID
    diag = precond_diag(marker).astype(dtype)
    x = np.zeros(marker.shape, dtype=dtype)
    r = div.astype(dtype).copy()
    z = np.where(fl, r / diag, 0).astype(dtype)
    p = z.copy()

    def dot(a, b):
        return t(np.sum((a * b)[fl], dtype=np.float64))

    rz_old = dot(r, z)
    r0 = dot(r, r)
    if r0 < 1e-30:
        return x, 0
    tol_sq = t(tol) ** 2 * r0
    for it in range(max_iter):
        Ap = apply_A(p, marker)
        pAp = dot(p, Ap)
        alpha = t(0) if abs(pAp) < 1e-30 else t(rz_old / pAp)
        x = np.where(fl, x + alpha * p, 0).astype(dtype)
        r = np.where(fl, r + (-alpha) * Ap, 0).astype(dtype)
        if (it + 1) % check_every == 0 or it == max_iter - 1:
            if dot(r, r) < tol_sq:
                return x, it + 1
        z = np.where(fl, r / diag, 0).astype(dtype)
        rz_new = dot(r, z)
        beta = t(0) if abs(rz_old) < 1e-30 else t(rz_new / rz_old)
        p = np.where(fl, z + beta * p, 0).astype(dtype)
        rz_old = rz_new
    return x, max_iter


def residual_norm(p, div, marker):
    """‖−div − A p‖₂ / ‖div‖₂ over fluid cells in FP64 (relative residual of
    the physically signed system)."""
    fl = marker == FLUID
    b = -div.astype(np.float64)
    r = (b - apply_A(p.astype(np.float64), marker))[fl]
    nb = np.linalg.norm(b[fl])
    return float(np.linalg.norm(r) / nb) if nb > 0 else 0.0


def subtract_pressure_grad(u, v, w, p, marker, dx, dt, rho):
    """u −= dt/(ρdx)·(p_r − p_l) on interior faces with ≥1 fluid and no solid
    neighbour; a non-fluid side reads p = 0 (solver3d.py:1364-1407)."""
    t = u.dtype.type
    scale = t(dt) / (t(rho) * t(dx))
    out = []
    for comp, f in ((0, u, ), (1, v), (2, w)):
        f = f.copy()
        n = marker.shape[comp]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[comp] = slice(0, n - 1)
        hi[comp] = slice(1, n)
        ml, mr = marker[tuple(lo)], marker[tuple(hi)]
        pl = np.where(ml == FLUID, p[tuple(lo)], 0).astype(f.dtype)
        pr = np.where(mr == FLUID, p[tuple(hi)], 0).astype(f.dtype)
        act = ((ml == FLUID) | (mr == FLUID)) & (ml != SOLID) & (mr != SOLID)
        face = [slice(None)] * 3
        face[comp] = slice(1, n)
        sub = f[tuple(face)]
        f[tuple(face)] = np.where(act, sub - scale * (pr - pl), sub).astype(f.dtype)
        out.append(f)
    return tuple(out)


# ------------------------------------------------------------------------ G2P
def sample3(f, fx, fy, fz):
    """Trilinear sample with coordinate and index clamping
    (gridmath.py:30-65); vectorised over points."""
    nmx, nmy, nmz = f.shape
    t = f.dtype.type
    fx = np.clip(fx, 0, nmx).astype(f.dtype)
    fy = np.clip(fy, 0, nmy).astype(f.dtype)
    fz = np.clip(fz, 0, nmz).astype(f.dtype)
    i0 = np.clip(np.floor(fx).astype(np.int64), 0, nmx - 1)
    j0 = np.clip(np.floor(fy).astype(np.int64), 0, nmy - 1)
    k0 = np.clip(np.floor(fz).astype(np.int64), 0, nmz - 1)
    i1 = np.minimum(i0 + 1, nmx - 1)
    j1 = np.minimum(j0 + 1, nmy - 1)
    k1 = np.minimum(k0 + 1, nmz - 1)
    sx = np.clip(fx - i0, 0, 1).astype(f.dtype)
    sy = np.clip(fy - j0, 0, 1).astype(f.dtype)
    sz = np.clip(fz - k0, 0, 1).astype(f.dtype)
    one = t(1)
    c00 = f[i0, j0, k0] * (one - sx) + f[i1, j0, k0] * sx
    c10 = f[i0, j1, k0] * (one - sx) + f[i1, j1, k0] * sx
    c01 = f[i0, j0, k1] * (one - sx) + f[i1, j0, k1] * sx
    c11 = f[i0, j1, k1] * (one - sx) + f[i1, j1, k1] * sx
    c0 = c00 * (one - sy) + c10 * sy
    c1 = c01 * (one - sy) + c11 * sy
    return c0 * (one - sz) + c1 * sz


def _lcoords(pos, dx, dtype):
    if dtype == np.float32:
        q = np.asarray(pos, dtype=np.float32) / np.float32(dx)
        return q, np.float32(0.5)
    return np.asarray(pos, dtype=np.float32).astype(np.float64) / float(np.float32(dx)), 0.5


def g2p_advect(pos, vel, u, v, w, us, vs, ws, dx, dt, flip_blend, dom):
    """FLIP/PIC gather + advection + clamp to [1.001dx, dom−1.001dx]
    (solver3d.py:1414-1463).  Returns (pos, vel) in the fields' precision."""
    dtype = u.dtype.type
    q, h = _lcoords(pos, dx, dtype)
    lx, ly, lz = q[:, 0], q[:, 1], q[:, 2]
    pic = np.stack([sample3(u, lx, ly - h, lz - h), sample3(v, lx - h, ly, lz - h),
                    sample3(w, lx - h, ly - h, lz)], axis=1)
    old = np.stack([sample3(us, lx, ly - h, lz - h), sample3(vs, lx - h, ly, lz - h),
                    sample3(ws, lx - h, ly - h, lz)], axis=1)
    fb = dtype(flip_blend)
    flip = np.asarray(vel).astype(dtype) + (pic - old)
    nvel = (fb * flip + (dtype(1) - fb) * pic).astype(dtype)
    p = np.asarray(pos, dtype=np.float32).astype(dtype)
    npos = p + nvel * dtype(dt)
    eps = dtype(np.float32(dx) * np.float32(1.001))
    for ax in range(3):
        npos[:, ax] = np.clip(npos[:, ax], eps, dtype(dom[ax]) - eps)
    return npos, nvel


def g2p_apic_advect(pos, u, v, w, dx, dt, dom):
    """APIC gather with clamped corner indices, affine matrix rows
    Σ w·u·(x_face − x_p)·3/dx², then advection + clamp (solver3d.py:1591-1732).
    Returns (pos, vel, C (n,3,3))."""
    dtype = u.dtype.type
    q, h = _lcoords(pos, dx, dtype)
    p = np.asarray(pos, dtype=np.float32).astype(dtype)
    dxc = dtype(np.float32(dx))
    inv = dtype(3.0) / (dxc * dxc)
    n = len(p)
    vel = np.zeros((n, 3), dtype=dtype)
    C = np.zeros((n, 3, 3), dtype=dtype)
    for comp, f in enumerate((u, v, w)):
        c = q.copy()
        for ax in range(3):
            if ax != comp:
                c[:, ax] = q[:, ax] - h
        i0, frac = _corners(c)
        hi = [f.shape[0] - 1, f.shape[1] - 1, f.shape[2] - 1]
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    idx = [np.clip(i0[:, 0] + di, 0, hi[0]), np.clip(i0[:, 1] + dj, 0, hi[1]),
                           np.clip(i0[:, 2] + dk, 0, hi[2])]
                    wx = frac[:, 0] if di else 1 - frac[:, 0]
                    wy = frac[:, 1] if dj else 1 - frac[:, 1]
                    wz = frac[:, 2] if dk else 1 - frac[:, 2]
                    wt = (wx * wy * wz).astype(dtype)
                    val = f[idx[0], idx[1], idx[2]]
                    vel[:, comp] += wt * val
                    fu = wt * val * inv
                    for ax in range(3):
                        off = 0.0 if ax == comp else 0.5
                        o = (idx[ax].astype(dtype) + dtype(off)) * dxc - p[:, ax]
                        C[:, comp, ax] += fu * o
    npos = p + vel * dtype(dt)
    eps = dtype(np.float32(dx) * np.float32(1.001))
    for ax in range(3):
        npos[:, ax] = np.clip(npos[:, ax], eps, dtype(dom[ax]) - eps)
    return npos, vel, C


# -------------------------------------------------------------------- substep
# ------------------------------------------------------------------ viscosity
def visc_sweep(u_old, u_in, r):
    """k3_jacobi_visc (solver3d.py:178-197): one Jacobi sweep of
    (I - r Lap) u = u_old on one component array; in-range neighbours summed
    in the order -x, +x, -y, +y, -z, +z; (u_old + r*sum) / (1 + r*count) in
    the arrays' precision, each operation rounded on its own."""
    t = u_in.dtype.type
    r = t(np.float32(r))
    s = np.zeros_like(u_in)
    cnt = np.zeros_like(u_in)
    one = t(1)
    for ax in range(3):
        for d in (-1, 1):
            nb = _shift(u_in, ax, d, 0)
            ok = _shift(np.ones_like(u_in), ax, d, 0)
            s = (s + nb).astype(u_in.dtype)
            cnt = (cnt + ok).astype(u_in.dtype)
    num = (np.asarray(u_old, u_in.dtype) + (r * s).astype(u_in.dtype)).astype(u_in.dtype)
    den = (one + (r * cnt).astype(u_in.dtype)).astype(u_in.dtype)
    return (num / den).astype(u_in.dtype)


def box_face_slices(box, c):
    """Index of a sub-dense box's faces in component c's face array (c None:
    cells); box = ((ox, oy, oz), (sx, sy, sz))."""
    lo, shape = box
    return tuple(slice(lo[a], lo[a] + shape[a] + (1 if a == c else 0)) for a in range(3))


def viscosity_stage(u, v, w, us, vs, ws, marker, r, iters, box=None):
    """The S2.13 stage of _step_impl (solver3d.py:3524-3544): right-hand side
    = the post-BC faces; each sweep's result becomes the component and the
    component before it moves into us/vs/ws (the reference swaps the arrays),
    so us ends as the last-but-one iterate (or unchanged for 0 sweeps); then
    the solid BC.  Returns (u, v, w, us, vs, ws).

    box: the sub-dense box of F3.7 (solver3d.py:3821-3906), on which the
    reference runs the same sweeps over its box-shaped arrays (the box edge
    is the stencil's array bound).  The sweeps then run on each component's
    box slice; outside the box the faces keep their pre-sweep values, in
    both the component and us/vs/ws (with no sweep at all, as without a
    box)."""
    if box is not None and iters > 0:
        cur = [u.copy(), v.copy(), w.copy()]
        sav = [u.copy(), v.copy(), w.copy()]
        sl = [box_face_slices(box, c) for c in range(3)]
        rhs = [cur[c][sl[c]].copy() for c in range(3)]
        bc = [cur[c][sl[c]].copy() for c in range(3)]
        bs = [sav[c][sl[c]].copy() for c in range(3)]
        for _ in range(iters):
            for c in range(3):
                new = visc_sweep(rhs[c], bc[c], r)
                bs[c], bc[c] = bc[c], new
        for c in range(3):
            cur[c][sl[c]] = bc[c]
            sav[c][sl[c]] = bs[c]
        cu, cv, cw = enforce_solid_bc(cur[0], cur[1], cur[2], marker)
        return cu, cv, cw, sav[0], sav[1], sav[2]
    rhs = (u.copy(), v.copy(), w.copy())
    cur = [u, v, w]
    sav = [us, vs, ws]
    for _ in range(iters):
        for c in range(3):
            new = visc_sweep(rhs[c], cur[c], r)
            sav[c], cur[c] = cur[c], new
    cu, cv, cw = enforce_solid_bc(cur[0], cur[1], cur[2], marker)
    return cu, cv, cw, sav[0], sav[1], sav[2]


# ------------------------------------------------------------ surface tension
def csf_indicator(marker, dtype=np.float32):
    """k3_indicator_from_marker (solver3d.py:212-229)."""
    return (np.asarray(marker) == FLUID).astype(dtype)


def box_blur(src):
    """k3_box_blur_centred (solver3d.py:235-256): mean over the in-range cells
    of the 3x3x3 block, summed with di outermost and dk innermost."""
    t = src.dtype
    s = np.zeros_like(src)
    n = np.zeros_like(src)
    ones = np.ones_like(src)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                v, ok = src, ones
                for ax, d in ((0, di), (1, dj), (2, dk)):
                    if d:
                        v = _shift(v, ax, d, 0)
                        ok = _shift(ok, ax, d, 0)
                s = (s + v).astype(t)
                n = (n + ok).astype(t)
    return (s / n).astype(t)


def _interior(shape):
    m = np.zeros(shape, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def csf_normal(chi, dx):
    """k3_csf_normal (solver3d.py:262-286): grad chi / (|grad chi| + 1e-8) by
    central differences over 2dx; zero on the outer cell layer."""
    t = chi.dtype.type
    h = t(2) * t(np.float32(dx))
    g = [((_shift(chi, ax, 1, 0) - _shift(chi, ax, -1, 0)) / h).astype(chi.dtype)
         for ax in range(3)]
    sq = ((g[0] * g[0]).astype(chi.dtype) + (g[1] * g[1]).astype(chi.dtype)).astype(chi.dtype)
    sq = (sq + (g[2] * g[2]).astype(chi.dtype)).astype(chi.dtype)
    mag = (np.sqrt(sq) + t(1e-8)).astype(chi.dtype)
    inside = _interior(chi.shape)
    return [np.where(inside, (gc / mag).astype(chi.dtype), t(0)) for gc in g]


def csf_curvature(n, dx):
    """k3_csf_curvature (solver3d.py:292-313): kappa = -div n by central
    differences; zero on the outer cell layer."""
    t = n[0].dtype.type
    h = t(2) * t(np.float32(dx))
    d = [((_shift(n[ax], ax, 1, 0) - _shift(n[ax], ax, -1, 0)) / h).astype(n[0].dtype)
         for ax in range(3)]
    k = -(((d[0] + d[1]).astype(n[0].dtype) + d[2]).astype(n[0].dtype))
    return np.where(_interior(n[0].shape), k, t(0)).astype(n[0].dtype)


def _face_masks(marker, comp):
    """Faces of component comp with both cells in range: (active = neither
    side solid, fluid = active and one side fluid), as face-shaped masks."""
    nx, ny, nz = marker.shape
    shape = [nx, ny, nz]
    shape[comp] += 1
    active = np.zeros(shape, dtype=bool)
    fluid = np.zeros(shape, dtype=bool)
    sl = [slice(None)] * 3
    sl[comp] = slice(1, shape[comp] - 1)
    lo = [slice(None)] * 3
    lo[comp] = slice(0, -1)
    hi = [slice(None)] * 3
    hi[comp] = slice(1, None)
    mlo, mhi = marker[tuple(lo)], marker[tuple(hi)]
    act = (mlo != SOLID) & (mhi != SOLID)
    active[tuple(sl)] = act
    fluid[tuple(sl)] = act & ((mlo == FLUID) | (mhi == FLUID))
    return active, fluid, tuple(lo), tuple(hi), tuple(sl)


def csf_apply(u, v, w, marker, dx, coeff, passes):
    """_apply_surface_tension (solver3d.py:3230-3300): indicator, `passes`
    blurs, normal, curvature; face += coeff*kappa_face*grad chi on faces with
    no solid side (k3_csf_apply_*, 316-385); then each component's mean change
    over faces next to fluid, summed exactly and rounded to the fields'
    precision, is subtracted there (k3_csf_sum_* / subtract_bias_*_dev,
    400-547).  Returns new (u, v, w)."""
    dt_ = u.dtype
    chi = csf_indicator(marker, dt_)
    for _ in range(passes):
        chi = box_blur(chi)
    kappa = csf_curvature(csf_normal(chi, dx), dx)
    out = []
    for comp, f in enumerate((u, v, w)):
        g = csf_face_apply(f, chi, kappa, marker, comp, dx, coeff)
        total, cnt = csf_face_sum(f, g, marker, comp)
        out.append(csf_face_subtract(g, marker, comp, total, cnt))
    return tuple(out)


def csf_face_apply(f, chi, kappa, marker, comp, dx, coeff):
    """k3_csf_apply_u/v/w (solver3d.py:316-385) on component comp: face +=
    coeff * kappa_face * grad(chi) on faces with no solid side."""
    dt_ = f.dtype
    t = dt_.type
    active, _, lo, hi, sl = _face_masks(marker, comp)
    grad = ((chi[hi] - chi[lo]) / t(np.float32(dx))).astype(dt_)
    kf = (t(0.5) * (kappa[hi] + kappa[lo]).astype(dt_)).astype(dt_)
    inc = ((t(np.float32(coeff)) * kf).astype(dt_) * grad).astype(dt_)
    g = f.copy()
    inner = g[sl]
    a = active[sl]
    inner[a] = (inner[a] + inc[a]).astype(dt_)
    g[sl] = inner
    return g


def csf_face_sum(before, after, marker, comp):
    """k3_csf_sum_* (solver3d.py:400-464): the change after - before summed
    exactly over the faces next to fluid (rounded to the fields' precision)
    and the number of those faces."""
    _, fluid, _, _, _ = _face_masks(marker, comp)
    delta = (after - before).astype(after.dtype)[fluid]
    total = delta.astype(np.float64).sum()
    if after.dtype == np.float32:
        total = np.float32(total)
    return total, int(fluid.sum())


def csf_face_subtract(g, marker, comp, total, cnt):
    """k3_csf_subtract_bias_*_dev (solver3d.py:478-547): faces next to fluid
    lose total / cnt when cnt > 0."""
    dt_ = g.dtype
    t = dt_.type
    g = g.copy()
    if cnt > 0:
        _, fluid, _, _, _ = _face_masks(marker, comp)
        mean = t(total) / t(cnt)
        g[fluid] = (g[fluid] - mean).astype(dt_)
    return g


# ------------------------------------------- per-particle attribute transfer
ATTR_COLOR, ATTR_SCALAR = 0, 1


def attr_p2g(pos, attr, dx, shape):
    """Trilinear scatter of per-particle values onto the cell-centred grid
    (k3_p2g_color solver3d.py:562-605, k3_p2g_scalar 670-707): fx = p/dx-0.5,
    corners in range only, w = (wx*wy)*wz and value*w in FP32.  The cell sums
    are formed in FP64 from those FP32 products (the order-free sum; the
    reference's FP32 atomics give it up to rounding in the order of arrival).
    attr (n,) or (n,c).  Returns (vsum (c,nx,ny,nz), wsum (nx,ny,nz)), FP64."""
    nx, ny, nz = shape
    a = np.asarray(attr, dtype=np.float32)
    a2 = a.reshape(len(a), -1)
    q = np.asarray(pos, dtype=np.float32) / np.float32(dx) - np.float32(0.5)
    i0 = np.floor(q).astype(np.int64)
    s = (q - i0.astype(np.float32)).astype(np.float32)
    one = np.float32(1)
    vsum = np.zeros((a2.shape[1], nx, ny, nz))
    wsum = np.zeros((nx, ny, nz))
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                ii, jj, kk = i0[:, 0] + di, i0[:, 1] + dj, i0[:, 2] + dk
                ok = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
                wx = s[:, 0] if di else one - s[:, 0]
                wy = s[:, 1] if dj else one - s[:, 1]
                wz = s[:, 2] if dk else one - s[:, 2]
                w = ((wx * wy).astype(np.float32) * wz).astype(np.float32)
                idx = (ii[ok], jj[ok], kk[ok])
                np.add.at(wsum, idx, w[ok].astype(np.float64))
                for c in range(a2.shape[1]):
                    np.add.at(vsum[c], idx, (a2[ok, c] * w[ok]).astype(np.float32))
    return vsum, wsum


def attr_normalize(vsum, wsum, mode):
    """k3_normalize_color (solver3d.py:607-631: value * (1/w)) or
    k3_normalize_scalar (710-727: value / w), FP32, 0 where w <= 1e-8.
    Returns (grid (c,nx,ny,nz) FP32, w FP32)."""
    w = np.asarray(wsum).astype(np.float32)
    v = np.asarray(vsum).astype(np.float32)
    ok = w > np.float32(1e-8)
    with np.errstate(divide="ignore", invalid="ignore"):
        if mode == ATTR_COLOR:
            g = v * (np.float32(1) / w)
        else:
            g = v / w
    return np.where(ok, g, np.float32(0)).astype(np.float32), w


def attr_g2p(pos, attr, grid, wgrid, dx):
    """k3_g2p_color (solver3d.py:633-664) / k3_g2p_scalar (729-755): FP32
    sample3 of each channel and of the weight at p/dx-0.5; a particle is
    updated only where the sampled weight exceeds 1e-6."""
    a = np.array(attr, dtype=np.float32)
    a2 = a.reshape(len(a), -1)
    q = np.asarray(pos, dtype=np.float32) / np.float32(dx) - np.float32(0.5)
    fx, fy, fz = q[:, 0], q[:, 1], q[:, 2]
    w = sample3(wgrid, fx, fy, fz)
    upd = w > np.float32(1e-6)
    for c in range(a2.shape[1]):
        a2[upd, c] = sample3(grid[c], fx, fy, fz)[upd]
    return a2.reshape(a.shape)


def attr_transfer(pos, attr, dx, shape, mode):
    """P2G -> normalize -> G2P of one attribute (_apply_color_transfer
    solver3d.py:3333-3359, _apply_scalar_transfer 3304-3330)."""
    vsum, wsum = attr_p2g(pos, attr, dx, shape)
    grid, w = attr_normalize(vsum, wsum, mode)
    return attr_g2p(pos, attr, grid, w, dx)


def attr_transfer_loop(pos, attr, dx, shape, mode):
    """Per-particle Python loop of the same transfer for tiny cases, written
    from the scalar kernels with explicit per-corner bounds checks."""
    nx, ny, nz = shape
    a = np.array(attr, dtype=np.float64).reshape(len(attr), -1)
    nc = a.shape[1]
    acc = {}
    for p in range(len(pos)):
        f = [float(np.float32(pos[p][d]) / np.float32(dx)) - 0.5 for d in range(3)]
        b = [int(np.floor(x)) for x in f]
        t = [f[d] - b[d] for d in range(3)]
        for corner in range(8):
            o = ((corner >> 2) & 1, (corner >> 1) & 1, corner & 1)
            c = tuple(b[d] + o[d] for d in range(3))
            if not all(0 <= c[d] < n for d, n in enumerate((nx, ny, nz))):
                continue
            w = 1.0
            for d in range(3):
                w *= t[d] if o[d] else 1.0 - t[d]
            e = acc.setdefault(c, [0.0] * (nc + 1))
            for ch in range(nc):
                e[ch] += a[p, ch] * w
            e[nc] += w
    grid = np.zeros((nc, nx, ny, nz))
    wg = np.zeros((nx, ny, nz))
    for c, e in acc.items():
        wg[c] = e[nc]
        if e[nc] > 1e-8:
            for ch in range(nc):
                grid[(ch,) + c] = e[ch] / e[nc]
    out = a.copy()
    for p in range(len(pos)):
        f = [float(np.float32(pos[p][d]) / np.float32(dx)) - 0.5 for d in range(3)]
        wv = float(sample3(wg, np.array([f[0]]), np.array([f[1]]), np.array([f[2]]))[0])
        if wv > 1e-6:
            for ch in range(nc):
                out[p, ch] = float(sample3(grid[ch], np.array([f[0]]), np.array([f[1]]),
                                           np.array([f[2]]))[0])
    return out.reshape(np.shape(attr))


def substep(state: dict, dt: float, pressure="pcg", iters=80, tol=1e-4, dtype=np.float64,
            gravity=-9.81, flip_blend=0.95, rho=1.0, transfer="flip", pcg_sign="physical",
            viscosity=0.0, viscosity_iters=12, surface_tension=0.0, csf_passes=2,
            visc_box=None):
    """One substep in the stage order of _step_impl (solver3d.py:3475-3667):
    clear → P2G → normalize → gravity → BC → [viscosity → BC] → [surface
    tension → BC] → divergence →
    pressure → grad → BC → G2P + advect → colour / temperature transfer.  ``state`` holds pos,
    vel, marker, dx, (APIC) affine and optionally color (n,3) / temperature (n,).
    pressure: "jacobi" / "gsrb" (fixed iterations), "pcg" (reference restatement;
    pcg_sign "physical" negates its result, "reference" keeps it) or
    "exact" (spsolve).  visc_box: the F3.7 sub-dense box the viscosity runs on
    (viscosity_stage).  Returns a new state plus the grid fields."""
    pos, vel, marker, dx = state["pos"], state["vel"], state["marker"], state["dx"]
    nx, ny, nz = marker.shape
    dom = [domain(n, dx) for n in (nx, ny, nz)]
    m = update_marker(marker, pos, dx)
    aff = state.get("affine") if transfer == "apic" else None
    if transfer == "apic" and aff is None:
        aff = np.zeros((len(pos), 3, 3), dtype=np.float32)
    u, v, w, uw, vw, ww = p2g(pos, vel, m, dx, dtype=dtype, affine=aff)
    u, us = normalize(u, uw)
    v, vs = normalize(v, vw)
    w, ws = normalize(w, ww)
    v = add_gravity(v, gravity, dt)
    u, v, w = enforce_solid_bc(u, v, w, m)
    if viscosity > 0.0:
        r = np.float32(dt * viscosity / (dx * dx))
        u, v, w, us, vs, ws = viscosity_stage(u, v, w, us, vs, ws, m, r, viscosity_iters,
                                              box=visc_box)
    if surface_tension > 0.0:
        coeff = np.float32(surface_tension * dt / max(rho, 1.0e-8))
        u, v, w = csf_apply(u, v, w, m, dx, coeff, max(1, int(csf_passes)))
        u, v, w = enforce_solid_bc(u, v, w, m)
    div = divergence(u, v, w, m, dx, dt, rho)
    info = {}
    if pressure == "jacobi":
        p = jacobi(div, m, iters, dtype=dtype)
    elif pressure == "gsrb":
        p = gsrb(div, m, iters, dtype=dtype)
    elif pressure == "pcg":
        x, it = pcg_reference(div, m, iters, tol=tol, dtype=dtype)
        p = -x if pcg_sign == "physical" else x
        info["iters"] = it
    elif pressure == "exact":
        p = solve_exact(div, m).astype(dtype)
    else:
        raise ValueError(f"unknown pressure_solver: {pressure!r}")
    u2, v2, w2 = subtract_pressure_grad(u, v, w, p, m, dx, dt, rho)
    u2, v2, w2 = enforce_solid_bc(u2, v2, w2, m)
    out = dict(state)
    out["marker"] = m
    if transfer == "apic":
        npos, nvel, C = g2p_apic_advect(pos, u2, v2, w2, dx, dt, dom)
        out["affine"] = C
    else:
        npos, nvel = g2p_advect(pos, vel, u2, v2, w2, us, vs, ws, dx, dt, flip_blend, dom)
    out["pos"], out["vel"] = npos, nvel
    p32 = np.asarray(npos, dtype=np.float32)
    if state.get("color") is not None:
        out["color"] = attr_transfer(p32, state["color"], dx, m.shape, ATTR_COLOR)
    if state.get("temperature") is not None:
        out["temperature"] = attr_transfer(p32, state["temperature"], dx, m.shape, ATTR_SCALAR)
    grid = dict(u=u2, v=v2, w=w2, us=us, vs=vs, ws=ws, div=div, p=p, marker=m, info=info)
    return out, grid


# -------------------------------------------------------------------- seeding
def seed_box_reference(lo, hi, ppc, dx):
    """Literal FlipSolver3D.seed_box positions (solver3d.py:2826-2910): cells
    [int(lo/dx), int(hi/dx)) per axis, per_axis = round(ppc^(1/3)), jitter
    0.3·(U−0.5)/per_axis from default_rng(0) drawn x,y,z per particle."""
    i0, i1 = int(lo[0] / dx), int(hi[0] / dx)
    j0, j1 = int(lo[1] / dx), int(hi[1] / dx)
    k0, k1 = int(lo[2] / dx), int(hi[2] / dx)
    per = int(round(ppc ** (1.0 / 3.0)))
    rng = np.random.default_rng(0)
    out = []
    for i in range(i0, i1):
        for j in range(j0, j1):
            for k in range(k0, k1):
                for sx in range(per):
                    for sy in range(per):
                        for sz in range(per):
                            px = (i + (sx + 0.5 + 0.3 * (rng.random() - 0.5)) / per) * dx
                            py = (j + (sy + 0.5 + 0.3 * (rng.random() - 0.5)) / per) * dx
                            pz = (k + (sz + 0.5 + 0.3 * (rng.random() - 0.5)) / per) * dx
                            out.append([px, py, pz])
    return np.array(out, dtype=np.float32).reshape(-1, 3)


def seed_cells_loop(cells, dx, seed, per_axis=2, jitter=0.25):
    """Pure-Python form of the configuration seeding recipe (used to pin the
    vectorised seeder in gpufluid_amd.seeding): per cell in C order, per
    sub-position (sx,sy,sz) in C order, one U(−jitter, jitter)·dx draw per axis
    in x,y,z order from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for (i, j, k) in cells:
        for sx in range(per_axis):
            for sy in range(per_axis):
                for sz in range(per_axis):
                    jx, jy, jz = rng.random(3)
                    out.append([(i + (sx + 0.5) / per_axis + (jx - 0.5) * 2 * jitter) * dx,
                                (j + (sy + 0.5) / per_axis + (jy - 0.5) * 2 * jitter) * dx,
                                (k + (sz + 0.5) / per_axis + (jz - 0.5) * 2 * jitter) * dx])
    return np.array(out, dtype=np.float32).reshape(-1, 3)
