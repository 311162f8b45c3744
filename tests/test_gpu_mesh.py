"""Mesh obstacles on the HIP path (gf_mesh_*: the inside test of
schemes/mesh_marker.py:39-80, 197-236 and the animated mesh obstacles of
prepare_frame, solver3d.py:2598-2631) against the oracle's brute force over
every triangle: identical solid cells and face velocities."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import meshgen as mg
from oracle import frame_oracle as fr

pytestmark = pytest.mark.gpu


def _mark(shape, dx, tris, marker=None):
    from gpufluid_amd.meshes import mark_mesh
    m = torch.zeros(shape, dtype=torch.int32, device="cuda:0") if marker is None else \
        torch.as_tensor(marker, device="cuda:0").contiguous()
    mark_mesh(m, tris, dx)
    return m.cpu().numpy()


def _cases():
    rng = np.random.default_rng(5)
    cube = mg.triangles(*mg.cube((0.213, 0.301, 0.155), (0.647, 0.588, 0.702)))
    rot = mg.triangles(*mg.cube((-0.2, -0.25, -0.18), (0.2, 0.25, 0.18))) @ mg.rotation(7).T + 0.5
    sphere = mg.triangles(*mg.icosphere((0.47, 0.42, 0.5), 0.31, level=2))
    big = mg.triangles(*mg.icosphere((0.5, 0.5, 0.5), 0.33, level=4))         # 5120 triangles
    tor = mg.triangles(*mg.torus((0.45, 0.4, 0.5), 0.25, 0.09, 48, 24))
    outside = mg.triangles(*mg.icosphere((0.9, 0.1, 0.5), 0.45, level=2))     # crosses the walls
    nan = np.concatenate([sphere, np.full((3, 3, 3), np.nan)])
    jitter = sphere + 1e-3 * rng.standard_normal(sphere.shape)
    return {"cube": cube, "rotated_cube": rot, "sphere": sphere, "reversed": sphere[:, ::-1],
            "sphere_5120": big, "torus": tor, "beyond_domain": outside, "with_nan": nan,
            "jittered": jitter, "empty": np.zeros((0, 3, 3))}


@pytest.mark.parametrize("name", list(_cases()))
def test_inside_cells_match_oracle(name):
    tris = _cases()[name]
    shape, dx = (40, 36, 44), 1.0 / 44
    print(f"[mesh] {name}: {len(tris)} triangles on {shape}", flush=True)
    got = _mark(shape, dx, tris) == 2
    ref = fr.mesh_inside(shape, dx, tris.astype(np.float32))
    np.testing.assert_array_equal(got, ref)
    if name not in ("empty",):
        assert got.any()


def test_faces_of_new_cells_and_existing_solids():
    """Cells already solid stay untouched (no face writes); new ones get
    the velocity on all six faces."""
    from gpufluid_amd.meshes import mark_mesh
    shape, dx = (24, 24, 24), 1.0 / 24
    walls = np.zeros(shape, np.int32)
    walls[[0, -1]] = 2
    walls[:, [0, -1]] = 2
    walls[:, :, [0, -1]] = 2
    sphere = [("sphere", (0.35, 0.5, 0.5), {"radius": 0.16}, (0.5, 0.25, -1.0))]
    base, su0, sv0, sw0 = fr.stamp_obstacles(walls, dx, sphere)
    tris = mg.triangles(*mg.cube((0.41, 0.33, 0.29), (0.77, 0.71, 0.68))).astype(np.float32)
    vel = (-2.0, 0.75, 3.0)
    dev = "cuda:0"
    m = torch.as_tensor(base, device=dev).contiguous()
    faces = [torch.as_tensor(a, device=dev).contiguous() for a in (su0, sv0, sw0)]
    mark_mesh(m, tris, dx, faces, vel)
    exp = fr.stamp_obstacles(walls, dx, sphere, [(tris, vel)])
    assert ((fr.mesh_inside(shape, dx, tris)) & (base == 2)).any()   # overlaps the sphere
    np.testing.assert_array_equal(m.cpu().numpy(), exp[0])
    for got, e in zip(faces, exp[1:]):
        np.testing.assert_array_equal(got.cpu().numpy(), e)


def test_reference_signatures(tmp_path):
    from gpufluid_amd.animation import cell_centers
    from gpufluid_amd.meshes import mark_solid_from_mesh_gpu, mesh_indicator_sdf_gpu
    shape, dx = (20, 16, 24), 1.0 / 24
    tris = mg.triangles(*mg.icosphere((0.4, 0.35, 0.5), 0.25, level=2)).astype(np.float32)
    cc = cell_centers(*shape, dx)
    m = torch.zeros(shape, dtype=torch.int32, device="cuda:0")
    mark_solid_from_mesh_gpu(m, cc, tris)
    ref = fr.mesh_inside(shape, dx, tris)
    np.testing.assert_array_equal(m.cpu().numpy() == 2, ref)
    sdf = mesh_indicator_sdf_gpu(cc, tris, dx)
    np.testing.assert_array_equal(sdf, np.where(ref, -dx, dx).astype(np.float32))
    with pytest.raises(ValueError):
        mark_solid_from_mesh_gpu(m, cc * 1.01 + 0.003, tris)


def test_prepare_frame_with_animated_mesh(tmp_path):
    """A moving mesh (OBJ file) after an analytic sphere, over three frames:
    marker and solid face velocities equal the oracle's stamping of the
    frame's obstacles, and the fluid steps around it."""
    import gfutil as gu
    from gpufluid_amd.animation import LinearMotion, KeyframeMotion, evaluate_center
    from gpufluid_amd.meshes import load_triangles
    v, f = mg.torus((0.0, 0.0, 0.0), 0.16, 0.06, 32, 16)
    path = tmp_path / "ring.obj"
    mg.write_obj(path, v, f)
    shape, dx, pos, vel = gu.dam_scene(n=32, seed=41)
    s = gu.make_solver(shape, dx, pos, vel)
    base_c = (0.5, 0.45, 0.5)
    mot = KeyframeMotion(keyframes=[(0, (0.45, 0.4, 0.5)), (4, (0.6, 0.5, 0.45))])
    s.add_animated_obstacle("sphere", (0.2, 0.7, 0.5), LinearMotion(velocity=(0.3, 0.0, 0.0)),
                            radius=0.1)
    s.add_animated_obstacle("mesh", base_c, mot, mesh_path=str(path), scale=1.25,
                            translate=(0.01, 0.0, -0.02))
    walls = np.zeros(shape, np.int32)
    walls[[0, -1]] = 2
    walls[:, [0, -1]] = 2
    walls[:, :, [0, -1]] = 2
    for frame in range(3):
        print(f"[mesh] prepare_frame {frame}", flush=True)
        s.prepare_frame(frame, 1.0 / 24)
        c_s = evaluate_center((0.2, 0.7, 0.5), LinearMotion(velocity=(0.3, 0.0, 0.0)), frame)
        c_m = evaluate_center(base_c, mot, frame)
        t_total = (np.asarray((0.01, 0.0, -0.02), np.float32)
                   + (np.asarray(c_m, np.float32) - np.asarray(base_c, np.float32)))
        tris = load_triangles(path, 1.25, t_total)
        v_m = s._obstacle_velocity_at(s._anim_specs[1], frame)
        exp = fr.stamp_obstacles(walls, dx, [("sphere", c_s, {"radius": 0.1}, (0.3, 0.0, 0.0))],
                                 [(tris, v_m)])
        np.testing.assert_array_equal(s.marker.numpy(), exp[0])
        for got, e in zip((s.solid_u, s.solid_v, s.solid_w), exp[1:]):
            np.testing.assert_array_equal(got.numpy(), e)
        assert (exp[0] == 2).sum() > (walls == 2).sum() + 100
        s.step(1.0 / 240, pressure_iters=100, pressure_solver="pcg")
        assert np.isfinite(s.get_particles()[0]).all()
