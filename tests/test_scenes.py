"""The seeded workload generators behind the benchmark configurations: the
1M block, the random-walk tilt of the 8M block and the batched ensemble."""
from __future__ import annotations

import math

import numpy as np

import fluid_amd as fa


def test_block_1m_is_the_configured_lattice():
    sc = fa.scenes.block_1m(seed=0, tracers=16)
    assert sc.n == 128 * 64 * 128
    assert math.isclose(sc.spacing, (0.4 / sc.n) ** (1.0 / 3.0), rel_tol=1e-12)
    assert math.isclose(sc.spacing, 7.2525e-3, rel_tol=1e-4)
    p = sc.pos4[:, :3]
    # on the floor, centred in x and z, inside the unit box
    assert p[:, 1].min() > -0.5 and p[:, 1].min() < -0.5 + sc.spacing
    assert abs(p[:, 0].mean()) < 1e-3 and abs(p[:, 2].mean()) < 1e-3
    assert np.abs(p).max() < 0.5
    assert sc.grid.dims == (71, 71, 71)


def test_gravity_is_a_unit_tilt_times_g():
    sc = fa.scenes.block_1m(seed=3, tracers=0)
    for t in (0.0, 0.3, 1.7):
        g = np.array(sc.gravity(t))
        assert math.isclose(np.linalg.norm(g), 9.81, rel_tol=1e-12)
        assert g[2] == 0.0
        assert abs(math.atan2(g[0], -g[1])) <= math.pi / 4 + 1e-12


def test_tilt_walk_is_seeded_with_the_configured_step():
    a = fa.scenes.tilt_walk(20001, 0.05, seed=7)
    b = fa.scenes.tilt_walk(20001, 0.05, seed=7)
    c = fa.scenes.tilt_walk(20001, 0.05, seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a[0] == 0.0
    steps = np.diff(a)
    assert abs(steps.std() - 0.05) < 0.002
    assert abs(steps.mean()) < 0.002


def test_walk_gravity_holds_each_frame_angle():
    sc = fa.scenes.small_block(seed=0)
    sc.walk = fa.scenes.tilt_walk(10, 0.05, seed=1)
    dt = 1.0 / 120.0
    for k in range(9):
        for frac in (0.01, 0.5, 0.99):
            g = sc.gravity((k + frac) * dt)
            assert math.isclose(math.atan2(g[0], -g[1]), sc.walk[k], rel_tol=0, abs_tol=1e-12)


def test_ensemble_counts_and_fills_stay_in_range():
    scenes = fa.scenes.ensemble(scenes=24, seed=5)
    counts = np.array([s.n for s in scenes])
    assert counts.min() >= 2048 * 0.9 and counts.max() <= 65536 * 1.1
    assert len({s.phase for s in scenes}) == len(scenes)
    again = fa.scenes.ensemble(scenes=24, seed=5)
    assert all(np.array_equal(x.pos4, y.pos4) for x, y in zip(scenes, again))
