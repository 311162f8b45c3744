"""Pins of the oracle's between-frame passes: reseed (sim/reseed.py:43-396)
and outflow compaction (solver3d.py:1946-1982, 2659-2730)."""
import numpy as np
import pytest

from oracle import frame_oracle as fr


def test_three_cell_hand_case():
    """Cells 0..2 along x (dx = 1): five particles in cell 0, none in the
    fluid-marked cell 1, one in cell 2.  min 2, max 3: cell 0 keeps its first
    three, cell 1 gets two, cell 2 one -> 3 emitted, 2 dropped."""
    marker = np.array([1, 1, 0], dtype=np.int32).reshape(3, 1, 1)
    pos = np.array([[0.1, .5, .5], [0.2, .5, .5], [2.5, .5, .5], [0.3, .5, .5],
                    [0.4, .5, .5], [0.9, .5, .5]], np.float32)
    alive, deficit, counts = fr.reseed_plan(pos, marker, 1.0, 2, 3, True)
    assert counts.ravel().tolist() == [5, 0, 1]
    assert deficit.ravel().tolist() == [0, 2, 1]
    assert alive.tolist() == [True, True, True, True, False, False]
    jit = np.zeros((3, 3), np.float32)
    e = fr.emit_positions(deficit, 1.0, jit)
    np.testing.assert_array_equal(e, [[1.5, .5, .5], [1.5, .5, .5], [2.5, .5, .5]])


def _scene(seed, n=3000, shape=(10, 8, 9)):
    rng = np.random.default_rng(seed)
    dx = 0.1
    ext = np.array(shape) * dx
    pos = (rng.random((n, 3)) * ext * np.array([0.5, 1, 1])).astype(np.float32)
    marker = np.zeros(shape, np.int32)
    marker[0] = marker[-1] = 2
    marker[:, 0] = marker[:, -1] = 2
    marker[7, 4, 4] = 2
    marker[6:8, 2:6, 2:6] = np.where(marker[6:8, 2:6, 2:6] == 2, 2, 1)
    return pos, marker, dx


@pytest.mark.parametrize("drop_solid", [True, False])
def test_reseed_invariants(drop_solid):
    pos, marker, dx = _scene(1)
    vel = np.random.default_rng(2).normal(size=pos.shape).astype(np.float32)
    mn, mx = 4, 6
    npos, nvel, _, n_emit, n_drop = fr.reseed(pos, vel, marker, dx, mn, mx, 0.3,
                                              np.random.default_rng(3), drop_solid)
    alive, deficit, counts = fr.reseed_plan(pos, marker, dx, mn, mx, drop_solid)
    n_alive = int(alive.sum())
    assert len(npos) == n_alive + n_emit and n_drop == len(pos) - n_alive
    # survivors keep their order and values
    np.testing.assert_array_equal(npos[:n_alive], pos[alive])
    np.testing.assert_array_equal(nvel[:n_alive], vel[alive])
    assert not nvel[n_alive:].any()
    # per-particle loop for the kept set
    _, key = fr.bin_cells(pos, dx, marker.shape)
    seen = {}
    for i, k in enumerate(key):
        if marker.reshape(-1)[k] == 2:
            assert alive[i] == (not drop_solid)
            continue
        seen[k] = seen.get(k, 0) + 1
        assert alive[i] == (seen[k] <= mx)
    # every target cell ends with at least min_per_cell particles, emitted
    # particles lie in their cells, non-solid cells hold at most max(mx, mn)
    _, k2 = fr.bin_cells(npos, dx, marker.shape)
    after = np.bincount(k2, minlength=marker.size)
    m = marker.reshape(-1)
    target = ((m == 1) | (counts.reshape(-1) > 0)) & (m != 2)
    assert (after[target] >= mn).all()
    assert (after[m != 2] <= max(mx, mn)).all()
    _, ke = fr.bin_cells(npos[n_alive:], dx, marker.shape)
    assert (np.bincount(ke, minlength=marker.size) == deficit.reshape(-1)).all()


def test_host_reference_counts_agree():
    """The reference host path's emission count is sum(clip(min - count)) over
    target cells and its cull count sum(clip(count - max)) over non-solid
    cells; the restatement's plan gives the same numbers."""
    pos, marker, dx = _scene(4)
    mn, mx = 3, 5
    alive, deficit, counts = fr.reseed_plan(pos, marker, dx, mn, mx, False)
    m = marker
    fluid = ((m == 1) | (counts > 0)) & (m != 2)
    assert deficit.sum() == np.clip(mn - counts, 0, None)[fluid].sum()
    over = np.clip(counts - mx, 0, None)
    over[m == 2] = 0
    assert (~alive).sum() == over.sum()


def test_outflow_mask_matches_loop():
    rng = np.random.default_rng(5)
    pos = rng.random((500, 3)).astype(np.float32)
    pos[:3] = [[0.2, 0.2, 0.2], [0.5, 0.5, 0.5], [0.2000001, 0.5, 0.3]]
    boxes = [((0.2, 0.2, 0.2), (0.5, 0.5, 0.5)), ((0.8, 0.0, 0.0), (1.0, 0.1, 1.0))]
    alive = fr.outflow_alive(pos, boxes)
    for i, p in enumerate(pos):
        inside = any(all(np.float32(lo[a]) <= p[a] <= np.float32(hi[a]) for a in range(3))
                     for lo, hi in boxes)
        assert alive[i] == (not inside)
    assert not alive[0] and not alive[1]


def test_obstacle_cells_against_fp64_distance_loop():
    shape, dx = (12, 10, 11), 0.1
    base = np.zeros(shape, np.int32)
    obs = [("sphere", (0.52, 0.47, 0.55), {"radius": 0.27}, (1.0, 0.0, 0.0)),
           ("box", (0.3, 0.2, 0.3), {"half_size": (0.12, 0.07, 0.21)}, (0.0, 2.0, 0.0)),
           ("cylinder_y", (0.8, 0.5, 0.5), {"radius": 0.16, "half_height": 0.31}, (0, 0, -1.5))]
    m, su, sv, sw = fr.stamp_obstacles(base, dx, obs)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                x, y, z = (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx
                ins = []
                (cx, cy, cz) = obs[0][1]
                ins.append((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= 0.27 ** 2)
                (cx, cy, cz), h = obs[1][1], obs[1][2]["half_size"]
                ins.append(abs(x - cx) <= h[0] and abs(y - cy) <= h[1] and abs(z - cz) <= h[2])
                (cx, cy, cz) = obs[2][1]
                ins.append((x - cx) ** 2 + (z - cz) ** 2 <= 0.16 ** 2 and abs(y - cy) <= 0.31)
                assert (m[i, j, k] == 2) == any(ins), (i, j, k)
    # a face takes the velocity of the last obstacle containing one of its cells
    cells = [fr.obstacle_sdf(fr.cell_centers(shape, dx), kd, c, p) <= 0 for kd, c, p, _ in obs]
    for i in range(shape[0] + 1):
        for j in range(shape[1]):
            for k in range(shape[2]):
                want = 0.0
                for b, (_, _, _, v) in enumerate(obs):
                    lo = i > 0 and cells[b][i - 1, j, k]
                    hi = i < shape[0] and cells[b][i, j, k]
                    if lo or hi:
                        want = v[0]
                assert su[i, j, k] == np.float32(want)
