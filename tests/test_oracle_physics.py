"""Pins the oracle's transfers and pressure projection with closed forms,
invariants and independent constructions (no reference import involved)."""
import numpy as np
import pytest

from oracle import flip_oracle as fo


def _rand_particles(rng, n, shape, dx, margin=1.0):
    lo = margin * dx
    hi = (np.array(shape) - margin) * dx
    return (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)


# --------------------------------------------------------------------- P2G
def test_p2g_vectorised_matches_per_particle_loop():
    rng = np.random.default_rng(0)
    shape, dx = (7, 6, 5), 1 / 8
    pos = _rand_particles(rng, 60, shape, dx, margin=0.0)
    pos[:4] = np.float32([[0, 0, 0], [7 / 8, 6 / 8, 5 / 8], [0.5 / 8, 3 / 8, 0.01], [0.3, 0.6, 0.2]])
    vel = rng.normal(size=(60, 3)).astype(np.float32)
    m = fo.wall_marker(*shape)
    a = fo.p2g(pos, vel, m, dx)
    b = fo.p2g_loop(pos, vel, m, dx)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-12)


def test_p2g_weights_partition_unity_for_interior_particle():
    shape, dx = (8, 8, 8), 1 / 8
    pos = np.float32([[0.41, 0.53, 0.37]])
    m = fo.wall_marker(*shape)
    _, _, _, uw, vw, ww = fo.p2g(pos, np.ones((1, 3), np.float32), m, dx)
    for wgt in (uw, vw, ww):
        assert abs(wgt.sum() - 1.0) < 1e-12
        assert (wgt > 0).sum() == 8


def test_p2g_uniform_velocity_is_reproduced_after_normalize():
    rng = np.random.default_rng(1)
    shape, dx = (10, 10, 10), 0.1
    pos = _rand_particles(rng, 4000, shape, dx)
    vel = np.tile(np.float32([0.3, -1.7, 2.5]), (len(pos), 1))
    m = fo.wall_marker(*shape)
    u, v, w, uw, vw, ww = fo.p2g(pos, vel, m, dx)
    for f, wgt, val in ((u, uw, 0.3), (v, vw, -1.7), (w, ww, 2.5)):
        nf, _ = fo.normalize(f, wgt)
        sel = wgt > 1e-8
        np.testing.assert_allclose(nf[sel], np.float32(val), rtol=1e-6)
        assert (nf[~sel] == 0).all()


def test_p2g_drops_out_of_grid_corners():
    shape, dx = (4, 4, 4), 0.25
    pos = np.float32([[0.01, 0.01, 0.01]])     # stencil reaches index -1 on v/w/u halves
    m = fo.wall_marker(*shape)
    _, _, _, uw, vw, ww = fo.p2g(pos, np.ones((1, 3), np.float32), m, dx)
    assert uw.sum() < 1.0 and vw.sum() < 1.0 and ww.sum() < 1.0


def test_apic_p2g_transfers_an_affine_field_exactly():
    """v_p + C (x_f - x_p) with v_p = a + G x_p, C = G is a + G x_f at every
    face: the normalised grid reproduces the affine field."""
    rng = np.random.default_rng(2)
    shape, dx = (8, 8, 8), 1 / 8
    pos = _rand_particles(rng, 3000, shape, dx)
    a = np.array([0.2, -0.4, 0.7])
    G = rng.normal(size=(3, 3))
    vel = (a + pos.astype(np.float64) @ G.T).astype(np.float32)
    C = np.tile(G.astype(np.float32), (len(pos), 1, 1))
    m = fo.wall_marker(*shape)
    fields = fo.p2g(pos, vel, m, dx, affine=C)
    for comp in range(3):
        f, wgt = fields[comp], fields[3 + comp]
        nf, _ = fo.normalize(f, wgt)
        idx = np.argwhere(wgt > 1e-3)
        xf = (idx + 0.5) * dx
        xf[:, comp] -= 0.5 * dx
        want = a[comp] + xf @ G[comp]
        np.testing.assert_allclose(nf[tuple(idx.T)], want, atol=2e-5)


# ----------------------------------------------------------- grid stencils
def test_gravity_bc_divergence_closed_form_at_rest():
    shape, dx, dt = (6, 7, 5), 0.1, 0.01
    m = fo.wall_marker(*shape)
    m[2, 3, 2] = fo.FLUID
    u = np.zeros((7, 7, 5)); v = np.zeros((6, 8, 5)); w = np.zeros((6, 7, 6))
    v = fo.add_gravity(v, -9.81, dt)
    assert np.allclose(v, -0.0981)
    u, v, w = fo.enforce_solid_bc(u, v, w, m)
    # only faces with no solid neighbour and off the domain edge keep g*dt
    assert v[2, 3, 2] == pytest.approx(-0.0981) and v[2, 4, 2] == pytest.approx(-0.0981)
    assert v[2, 1, 2] == 0 and v[2, 6, 2] == 0 and v[0, 3, 2] == 0
    div = fo.divergence(u, v, w, m, dx, dt, 1.0)
    assert (div[m != fo.FLUID] == 0).all()
    assert div[2, 3, 2] == pytest.approx(0.0)


def test_divergence_of_linear_field():
    shape, dx, dt, rho = (5, 5, 5), 0.2, 0.05, 2.0
    m = np.full(shape, fo.FLUID, np.int32)
    i = np.arange(6.0)
    u = np.broadcast_to((1.5 * i)[:, None, None], (6, 5, 5)).copy()
    v = np.broadcast_to((-0.5 * i)[None, :, None], (5, 6, 5)).copy()
    w = np.broadcast_to((0.25 * i)[None, None, :], (5, 5, 6)).copy()
    div = fo.divergence(u, v, w, m, dx, dt, rho)
    np.testing.assert_allclose(div, (1.5 - 0.5 + 0.25) * rho * dx / dt)


# ----------------------------------------------------------------- pressure
def _blob_marker(shape, seed):
    rng = np.random.default_rng(seed)
    m = fo.wall_marker(*shape)
    inner = rng.random(shape) < 0.6
    inner[0], inner[-1], inner[:, 0], inner[:, -1], inner[:, :, 0], inner[:, :, -1] = [False] * 6
    m[inner] = fo.FLUID
    return m


def test_apply_A_equals_assembled_matrix():
    m = _blob_marker((9, 8, 7), 3)
    A, fluid, _ = fo.poisson_matrix(m)
    rng = np.random.default_rng(4)
    x = np.zeros(m.shape)
    x[tuple(fluid.T)] = rng.normal(size=len(fluid))
    y = fo.apply_A(x, m)
    np.testing.assert_allclose(y[tuple(fluid.T)], A @ x[tuple(fluid.T)], atol=1e-12)
    assert abs(A - A.T).max() == 0
    # diagonally dominant: diag - sum |offdiag| = number of air neighbours >= 0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - d
    assert (d - off >= 0).all()


def test_one_cell_and_sign_convention():
    """Worked example: one fluid cell with air on both sides along x."""
    m = np.full((3, 1, 1), fo.AIR, np.int32)
    m[1, 0, 0] = fo.FLUID
    div = np.zeros((3, 1, 1))
    div[1, 0, 0] = 4.0
    p = fo.jacobi(div, m, 5, dtype=np.float64)
    assert p[1, 0, 0] == pytest.approx(-2.0)       # p = -div/diag, diag = 2
    x, it = fo.pcg_reference(div, m, 10, tol=1e-12, dtype=np.float64)
    assert x[1, 0, 0] == pytest.approx(+2.0)       # the reference PCG returns -p
    assert fo.solve_exact(div, m)[1, 0, 0] == pytest.approx(-2.0)


def test_tridiagonal_column_matches_dense_solve():
    n = 12
    m = np.full((n, 1, 1), fo.FLUID, np.int32)
    m[0] = fo.SOLID
    m[-1] = fo.AIR
    rng = np.random.default_rng(5)
    div = np.zeros((n, 1, 1))
    div[1:-1, 0, 0] = rng.normal(size=n - 2)
    k = n - 2
    # diag counts non-solid neighbours: 1 for the cell on the solid floor,
    # 2 elsewhere (the top cell's air neighbour is a Dirichlet p = 0 side)
    T = 2.0 * np.eye(k) - np.eye(k, k=1) - np.eye(k, k=-1)
    T[0, 0] = 1.0
    want = np.linalg.solve(T, -div[1:-1, 0, 0])
    got = fo.solve_exact(div, m)[1:-1, 0, 0]
    np.testing.assert_allclose(got, want, rtol=1e-12)
    pj = fo.jacobi(div, m, 4000, dtype=np.float64)[1:-1, 0, 0]
    np.testing.assert_allclose(pj, want, rtol=1e-6, atol=1e-9)


def test_jacobi_fixed_point_and_pcg_reference_match_direct_solve():
    m = _blob_marker((10, 9, 8), 6)
    rng = np.random.default_rng(7)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0.0)
    exact = fo.solve_exact(div, m)
    assert fo.residual_norm(exact, div, m) < 1e-12
    x, it = fo.pcg_reference(div, m, 500, tol=1e-12, dtype=np.float64)
    np.testing.assert_allclose(-x, exact, atol=1e-9 * np.abs(exact).max())
    assert it % 10 == 0 or it == 500                # checked every 10 iterations
    # one Jacobi sweep from the exact solution leaves it unchanged
    np.testing.assert_allclose(fo.jacobi_sweep(exact, div, m), exact, atol=1e-12)


def test_jacobi_fp32_sum_order_is_reference_order():
    """FP32 Jacobi equals an element-wise loop in the -x,+x,-y,+y,-z,+z order."""
    m = _blob_marker((6, 6, 6), 8)
    rng = np.random.default_rng(9)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    p_in = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    got = fo.jacobi_sweep(p_in, div, m)
    want = np.zeros_like(p_in)
    n = m.shape
    for i, j, k in np.argwhere(m == fo.FLUID):
        s = np.float32(0)
        d = np.float32(0)
        for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
            a, b, c = i + di, j + dj, k + dk
            if 0 <= a < n[0] and 0 <= b < n[1] and 0 <= c < n[2] and m[a, b, c] != fo.SOLID:
                d = np.float32(d + 1)
                if m[a, b, c] == fo.FLUID:
                    s = np.float32(s + p_in[a, b, c])
        want[i, j, k] = 0 if d < 0.5 else np.float32(np.float32(s - div[i, j, k]) / d)
    np.testing.assert_array_equal(got, want)


def test_projection_makes_fluid_cells_divergence_free():
    rng = np.random.default_rng(10)
    shape, dx, dt = (12, 12, 12), 1 / 12, 1 / 240
    cells = np.argwhere(np.ones((4, 6, 10), bool)) + 1
    pos = ((cells[:, None, :] + rng.random((len(cells), 8, 3))) * dx).reshape(-1, 3).astype(np.float32)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    st = {"pos": pos, "vel": vel, "marker": fo.wall_marker(*shape), "dx": dx}
    _, grid = fo.substep(st, dt, pressure="exact")
    m = grid["marker"]
    d = fo.divergence(grid["u"], grid["v"], grid["w"], m, dx, dt, 1.0)
    d0 = np.abs(grid["div"]).max()
    assert np.abs(d[m == fo.FLUID]).max() < 1e-9 * d0


# --------------------------------------------------------------------- G2P
def test_sample3_reproduces_trilinear_fields():
    rng = np.random.default_rng(11)
    f = np.fromfunction(lambda i, j, k: 0.5 + 1.5 * i - 2.0 * j + 0.25 * k + 0.1 * i * j * k,
                        (6, 7, 5))
    pts = rng.random((100, 3)) * np.array([5, 6, 4])
    got = fo.sample3(f, pts[:, 0], pts[:, 1], pts[:, 2])
    x, y, z = pts.T
    want = 0.5 + 1.5 * x - 2.0 * y + 0.25 * z + 0.1 * x * y * z
    np.testing.assert_allclose(got, want, rtol=1e-12)
    # clamping: outside points take the boundary value (constant extrapolation)
    out = fo.sample3(f, np.array([-3.0, 10.0]), np.array([0.0, 6.0]), np.array([0.0, 4.0]))
    np.testing.assert_allclose(out, [f[0, 0, 0], f[5, 6, 4]])


def test_pic_equals_flip_with_zero_blend_and_clamp_box():
    rng = np.random.default_rng(12)
    shape, dx, dt = (8, 8, 8), 1 / 8, 0.5
    m = fo.wall_marker(*shape)
    u = rng.normal(size=(9, 8, 8)); v = rng.normal(size=(8, 9, 8)); w = rng.normal(size=(8, 8, 9))
    us, vs, ws = u * 0.3, v * 0.1, w * 2
    pos = _rand_particles(rng, 200, shape, dx)
    vel = rng.normal(size=pos.shape).astype(np.float32) * 40
    dom = [fo.domain(8, dx)] * 3
    p0, v0 = fo.g2p_advect(pos, vel, u, v, w, us, vs, ws, dx, dt, 0.0, dom)
    p1, v1 = fo.g2p_advect(pos, vel * 0 + 7, u, v, w, us * 0, vs * 0, ws * 0, dx, dt, 0.0, dom)
    np.testing.assert_array_equal(v0, v1)     # PIC ignores particle velocity and old grid
    eps = np.float32(dx) * np.float32(1.001)
    assert p0.min() >= eps - 1e-12 and p0.max() <= dom[0] - eps + 1e-12


def test_apic_g2p_affine_reconstruction_closed_form():
    """For u = a*x the reconstructed C row is (3*a*s(1-s), 0, 0) with s the
    particle's fractional offset in the u-face lattice (D^-1 = 3/dx^2)."""
    shape, dx = (8, 8, 8), 1 / 8
    a = 1.7
    u = np.fromfunction(lambda i, j, k: a * i * dx, (9, 8, 8))
    v = np.zeros((8, 9, 8)); w = np.zeros((8, 8, 9))
    s = np.array([0.1, 0.25, 0.5, 0.8])
    pos = np.stack([(3 + s) * dx, np.full(4, 4.5 * dx), np.full(4, 3.5 * dx)], axis=1).astype(np.float32)
    dom = [fo.domain(8, dx)] * 3
    _, vel, C = fo.g2p_apic_advect(pos, u, v, w, dx, 0.0, dom)
    sx = pos[:, 0].astype(np.float64) / np.float32(dx) - 3
    np.testing.assert_allclose(vel[:, 0], a * pos[:, 0].astype(np.float64), rtol=1e-6)
    np.testing.assert_allclose(C[:, 0, 0], 3 * a * sx * (1 - sx), rtol=1e-5)
    np.testing.assert_allclose(C[:, 0, 1:], 0, atol=1e-9)


# ------------------------------------------------------------------ GS-RB
def test_gsrb_half_sweep_matches_sequential_in_place_loop():
    """The colour-at-once GS-RB half sweep equals the reference's in-place
    loop over the cells of one colour (solver3d.py:1204-1267), visiting cells
    in any order, in FP32 with the -x,+x,-y,+y,-z,+z sum order."""
    m = _blob_marker((6, 7, 5), 4)
    rng = np.random.default_rng(11)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    p0 = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    n = m.shape
    for color in (0, 1):
        got = fo.gsrb_half_sweep(p0, div, m, color)
        want = p0.copy()
        cells = [tuple(c) for c in np.argwhere(np.ones(n, bool))]
        rng.shuffle(cells)                          # in-place, any visiting order
        for i, j, k in cells:
            if (i + j + k) % 2 != color:
                continue
            if m[i, j, k] != fo.FLUID:
                want[i, j, k] = 0
                continue
            s = np.float32(0)
            d = np.float32(0)
            for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                a, b, c = i + di, j + dj, k + dk
                if 0 <= a < n[0] and 0 <= b < n[1] and 0 <= c < n[2] and m[a, b, c] != fo.SOLID:
                    d = np.float32(d + 1)
                    if m[a, b, c] == fo.FLUID:
                        s = np.float32(s + want[a, b, c])
            want[i, j, k] = 0 if d < 0.5 else np.float32(np.float32(s - div[i, j, k]) / d)
        np.testing.assert_array_equal(got, want)


def test_gsrb_converges_to_the_direct_solve_faster_than_jacobi():
    m = _blob_marker((10, 9, 8), 6)
    rng = np.random.default_rng(12)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0.0)
    exact = fo.solve_exact(div, m)
    scale = np.abs(exact).max()
    assert np.abs(fo.gsrb(div, m, 300, dtype=np.float64) - exact).max() < 1e-9 * scale
    err_gs = np.abs(fo.gsrb(div, m, 20, dtype=np.float64) - exact).max() / scale
    err_j = np.abs(fo.jacobi(div, m, 20, dtype=np.float64) - exact).max() / scale
    assert err_gs < 0.5 * err_j
    # one red + black sweep from the exact solution leaves it unchanged
    p = fo.gsrb_half_sweep(exact, div, m, 0)
    np.testing.assert_allclose(fo.gsrb_half_sweep(p, div, m, 1), exact, atol=1e-12)


def _linear_face_fields(coef, shape, dx):
    """Face arrays of the linear functions f_c(x) = coef[c,0] + coef[c,1:]·x
    sampled at the MAC face positions: u at (i, j+1/2, k+1/2)·dx, v at
    (i+1/2, j, k+1/2)·dx, w at (i+1/2, j+1/2, k)·dx."""
    nx, ny, nz = shape
    out = []
    for c, (sh, off) in enumerate((((nx + 1, ny, nz), (0.0, 0.5, 0.5)),
                                   ((nx, ny + 1, nz), (0.5, 0.0, 0.5)),
                                   ((nx, ny, nz + 1), (0.5, 0.5, 0.0)))):
        i, j, k = np.meshgrid(*(np.arange(s, dtype=np.float64) for s in sh), indexing="ij")
        x = np.stack([(i + off[0]) * dx, (j + off[1]) * dx, (k + off[2]) * dx])
        out.append(coef[c, 0] + np.tensordot(coef[c, 1:], x, axes=1))
    return out


@pytest.mark.parametrize("fb", [0.5, 0.95, 1.0])
def test_flip_blend_closed_form(fb):
    """The FLIP branch of g2p_advect against the reference formula written
    out by hand (solver3d.py:1447-1453): with grid fields u (post-projection)
    and us (pre-projection), v_new = fb*(v_p + u(x_p) - us(x_p)) +
    (1-fb)*u(x_p), x_new = x_p + v_new*dt.  Constant fields u = a, us = a - c
    give v_new = fb*(v_p + c) + (1-fb)*a; linear fields pin the half-cell
    offsets of each component's sampling lattice.  A swapped u/us, a sign
    error in (pic - old), a transposed component or a wrong stagger fails."""
    rng = np.random.default_rng(21)
    shape, dx, dt = (10, 12, 9), 1 / 12, 1e-3
    dom = [fo.domain(n, dx) for n in shape]
    n = 300
    # particles at least one cell inside the domain, so no sample is clamped
    lo, hi = 1.5 * dx, (np.array(shape) - 1.5) * dx
    pos = (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    x = pos.astype(np.float64)
    # constant fields, one value per component
    a = np.array([0.7, -1.3, 2.1])
    c = np.array([-0.4, 0.9, 0.25])
    u, v, w = (np.full(s, a[i]) for i, s in enumerate(((11, 12, 9), (10, 13, 9), (10, 12, 10))))
    us, vs, ws = u - c[0], v - c[1], w - c[2]
    p1, v1 = fo.g2p_advect(pos, vel, u, v, w, us, vs, ws, dx, dt, fb, dom)
    want = fb * (vel.astype(np.float64) + c) + (1 - fb) * a
    np.testing.assert_allclose(v1, want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(p1, x + want * dt, rtol=1e-12, atol=1e-12)
    # linear fields: pic and old are the functions at the particle position
    cu = rng.normal(size=(3, 4))
    cs = rng.normal(size=(3, 4))
    dx32 = float(np.float32(dx))       # the lattice spacing the FP32 position divides by
    u, v, w = _linear_face_fields(cu, shape, dx32)
    us, vs, ws = _linear_face_fields(cs, shape, dx32)
    p2, v2 = fo.g2p_advect(pos, vel, u, v, w, us, vs, ws, dx, dt, fb, dom)
    pic = cu[:, 0] + x @ cu[:, 1:].T
    old = cs[:, 0] + x @ cs[:, 1:].T
    want = fb * (vel.astype(np.float64) + pic - old) + (1 - fb) * pic
    np.testing.assert_allclose(v2, want, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(p2, x + want * dt, rtol=1e-9, atol=1e-9)
