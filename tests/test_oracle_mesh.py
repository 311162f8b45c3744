"""Pins of the oracle's mesh inside test (k3_mesh_to_solid,
schemes/mesh_marker.py:39-80; mesh obstacles of prepare_frame,
solver3d.py:2598-2631) and of the OBJ / STL loader (trimesh's triangles,
solver3d.py:2607-2613)."""
import numpy as np
import pytest

import meshgen as mg
from oracle import frame_oracle as fr

SHAPE = (20, 18, 22)
DX = 1.0 / 22


def _centres():
    return fr.cell_centers(SHAPE, DX).astype(np.float64)


def test_closed_cube_gives_the_box():
    """Box faces off the cell centres: inside cells are exactly those whose
    centre lies in the box (an odd number of crossings only from inside)."""
    lo, hi = (0.213, 0.301, 0.155), (0.647, 0.588, 0.702)
    tris = mg.triangles(*mg.cube(lo, hi))
    got = fr.mesh_inside(SHAPE, DX, tris)
    g = _centres()
    ref = np.all((g > np.array(lo)) & (g < np.array(hi)), axis=-1)
    assert got.sum() > 100
    np.testing.assert_array_equal(got, ref)


def test_reversed_winding_changes_nothing():
    v, f = mg.icosphere((0.45, 0.4, 0.5), 0.3, level=1)
    tris = mg.triangles(v, f)
    np.testing.assert_array_equal(fr.mesh_inside(SHAPE, DX, tris),
                                  fr.mesh_inside(SHAPE, DX, tris[:, ::-1]))


def test_polyhedral_sphere_between_its_radii():
    """Cells nearer than the polyhedron's inscribed radius are inside, cells
    beyond the circumscribed radius outside."""
    c, r = np.array([0.47, 0.42, 0.5]), 0.31
    v, f = mg.icosphere(c, r, level=2)
    tris = mg.triangles(v, f)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    r_in = np.abs(np.einsum("ij,ij->i", n, tris[:, 0] - c)).min()
    got = fr.mesh_inside(SHAPE, DX, tris)
    d = np.linalg.norm(_centres() - c, axis=-1)
    assert got[d < r_in - 1e-6].all()
    assert not got[d > r + 1e-6].any()
    assert 0.9 * (4 / 3 * np.pi * r_in ** 3) < got.sum() * DX ** 3 < 1.1 * (4 / 3 * np.pi * r ** 3)


def test_torus_is_not_convex():
    """Rays through the hole cross four times: the hole stays outside."""
    c, R, r = np.array([0.45, 0.4, 0.5]), 0.25, 0.09
    nu, nv = 48, 24
    tris = mg.triangles(*mg.torus(c, R, r, nu, nv))
    got = fr.mesh_inside(SHAPE, DX, tris)
    g = _centres() - c
    ring = np.sqrt(g[..., 0] ** 2 + g[..., 2] ** 2)
    sdf = np.sqrt((ring - R) ** 2 + g[..., 1] ** 2) - r
    tol = (R + r) * (1 - np.cos(np.pi / nu)) + r * (1 - np.cos(np.pi / nv)) + 1e-6
    assert got[sdf < -tol].all()
    assert not got[sdf > 1e-6].any()
    assert not got[ring < R - r - 1e-6].any()       # the hole


def test_vectorised_matches_scalar_loop():
    shape, dx = (6, 5, 7), 1.0 / 7
    v, f = mg.cube((0.1, 0.12, 0.2), (0.7, 0.6, 0.8))
    tris = mg.triangles(v, f) @ mg.rotation(3).T * 0.8 + 0.45
    tris = np.concatenate([tris, mg.triangles(*mg.icosphere((0.4, 0.3, 0.5), 0.2, 0))])
    np.testing.assert_array_equal(fr.mesh_inside(shape, dx, tris),
                                  fr.mesh_inside_loop(shape, dx, tris))


def test_degenerate_and_non_finite_triangles_never_cross():
    shape = (6, 6, 6)
    flat = np.array([[[0.0, 0.2, 0.2], [1.0, 0.2, 0.2], [0.5, 0.2, 0.2]]])  # no y-z area
    nan = np.array([[[np.nan, 0.1, 0.1], [0.9, 0.9, 0.1], [0.9, 0.1, 0.9]]])
    assert not fr.mesh_inside(shape, 1 / 6, flat).any()
    assert not fr.mesh_inside(shape, 1 / 6, nan).any()


def test_stamp_meshes_after_analytic_obstacles():
    """A mesh marks only cells not yet solid and writes its velocity into
    their faces, after (and over) the analytic obstacles' faces."""
    shape, dx = (16, 16, 16), 1 / 16
    base = fr.stamp_obstacles(np.zeros(shape, np.int32), dx, [])[0]
    sphere = ("sphere", (0.3, 0.5, 0.5), {"radius": 0.15}, (1.0, 0.0, 0.0))
    box = mg.triangles(*mg.cube((0.33, 0.31, 0.32), (0.71, 0.69, 0.68)))
    m0, su0, _, _ = fr.stamp_obstacles(base, dx, [sphere])
    m, su, sv, sw = fr.stamp_obstacles(base, dx, [sphere], [(box, (0.0, -2.0, 0.5))])
    inside_box = fr.mesh_inside(shape, dx, box)
    new = inside_box & (m0 != 2)
    assert new.any() and (inside_box & (m0 == 2)).any()
    np.testing.assert_array_equal(m, np.where(new, 2, m0))
    assert (sv[:, :-1][new] == -2.0).all() and (sw[:, :, 1:][new] == 0.5).all()
    assert (su[:-1][new] == 0.0).all()
    # faces of sphere cells away from the box keep the sphere's velocity
    keep = (m0 == 2) & ~np.pad(new, 1)[:-2, 1:-1, 1:-1] & ~new & (base != 2)
    assert (su[:-1][keep] == su0[:-1][keep]).all() and (su[:-1][keep] == 1.0).any()


@pytest.mark.parametrize("fmt", ["obj", "stl_binary", "stl_ascii"])
def test_loader_reads_the_triangles(tmp_path, fmt):
    from gpufluid_amd.meshes import load_triangles
    v, f = mg.icosphere((0.1, 0.2, 0.3), 0.7, level=1)
    v = v + 1e-3 * np.random.default_rng(1).random(v.shape)
    tris = mg.triangles(v, f)
    path = tmp_path / ("m.obj" if fmt == "obj" else "m.stl")
    if fmt == "obj":
        mg.write_obj(path, v, f)
        ref = tris
    elif fmt == "stl_binary":
        mg.write_stl_binary(path, tris)
        ref = tris.astype(np.float32).astype(np.float64)
    else:
        mg.write_stl_ascii(path, tris)
        ref = tris
    got = load_triangles(path, scale=1.5, translate=(0.25, -0.5, 0.125))
    exp = (ref * 1.5 + np.array([0.25, -0.5, 0.125], np.float32).astype(np.float64)).astype(np.float32)
    assert got.dtype == np.float32 and got.shape == (len(f), 3, 3)
    np.testing.assert_array_equal(got, exp)


def test_loader_fans_polygons(tmp_path):
    from gpufluid_amd.meshes import load_triangles
    v, quads = mg.cube((0, 0, 0), (1, 2, 3))
    path = tmp_path / "box.obj"
    mg.write_obj(path, v, quads)
    got = load_triangles(path)
    np.testing.assert_array_equal(got, mg.triangles(v, quads).astype(np.float32))
    assert len(got) == 12
    with pytest.raises(ValueError):
        load_triangles(tmp_path / "box.ply")
