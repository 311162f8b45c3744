"""Host CFL substep count (S2.10, solver3d.py:2218-2238) against the
reference's known answers (tests/test_s2_6_3_pcg.py:11-20) and its closed
form."""
import math

import numpy as np

from gpufluid_amd.solver3d import cfl_substep_count


def test_static_particles_take_one_substep():
    assert cfl_substep_count(np.zeros((100, 3), np.float32), dx=0.01, target_dt=0.01) == 1
    assert cfl_substep_count(np.zeros((0, 3), np.float32), dx=0.01, target_dt=0.01) == 1


def test_fast_particle_is_clamped():
    vel = np.array([[0.0, 100.0, 0.0]], np.float32)
    # cfl*dx/v = 5e-5 s -> 800 substeps, clamped to 32
    assert cfl_substep_count(vel, dx=0.01, target_dt=0.04, cfl=0.5, max_substeps=32) == 32


def test_closed_form_below_the_clamp():
    vel = np.array([[3.0, 4.0, 0.0], [0.1, 0.0, 0.0]], np.float32)   # |v|max = 5
    n = cfl_substep_count(vel, dx=0.02, target_dt=0.01, cfl=0.5, max_substeps=64)
    assert n == math.ceil(0.01 / (0.5 * 0.02 / 5.0)) == 5
