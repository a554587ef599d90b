"""The library's host helpers against the reference's own CPU tests.

Mirrors crates/fluid-core/src/sim.rs:358-567 (kernel, lattice, grid,
seeding, byte layouts), render.rs:3695-3728 (substep schedule) and
lib.rs:43-72 (body force). No device is needed: these are host functions
of libfluid_sim.so.
"""
import math

import numpy as np
import pytest

import fluid_amd as fa

SLAB_DEPTH = 4.0 * 0.00765
EXTENT = (0.0356, 0.077)
HALF = (EXTENT[0], EXTENT[1], SLAB_DEPTH / 2)
REST_DENSITY = 998.2


def test_the_kernel_integrates_to_one():  # sim.rs:358-373
    h = 0.004
    step = h / 20.0
    n = int(2.0 * h / step) + 1
    ax = np.arange(-n, n + 1)
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    r = np.sqrt((x * x + y * y + z * z).astype(np.float32)) * np.float32(step)
    # the library's kernel at every distinct radius, then summed with multiplicity
    radii, inv = np.unique(r.ravel(), return_inverse=True)
    w = np.array([fa.kernel(float(rr), h) for rr in radii])
    total = float(np.sum(w[inv]) * step ** 3)
    assert abs(total - 1.0) < 0.01, total


def test_a_lattice_at_spacing_d_recovers_rest_density():  # sim.rs:375-393
    d = 0.002
    h = 1.2 * d
    mass = np.float32(REST_DENSITY) * np.float32(d) ** 3
    rho = 0.0
    for ix in range(-3, 4):
        for iy in range(-3, 4):
            for iz in range(-3, 4):
                r = math.sqrt(ix * ix + iy * iy + iz * iz) * d
                rho += mass * fa.kernel(r, h)
    assert abs(rho - REST_DENSITY) / REST_DENSITY < 0.02


def test_every_seeded_particle_lands_in_a_valid_cell():  # sim.rs:395-405
    grid = fa.Grid.new(HALF, 0.0048)
    pts = fa.seed_slab(0.002, HALF, 0.5)
    assert len(pts) > 0
    for p in pts[:: max(1, len(pts) // 500)]:
        assert abs(p[0]) <= EXTENT[0] and abs(p[1]) <= EXTENT[1] and abs(p[2]) <= SLAB_DEPTH * 0.5
        assert grid.cell_of(p[:3]) < grid.cell_count()


def test_neighbouring_positions_map_to_adjacent_cells():  # sim.rs:407-415
    grid = fa.Grid.new(HALF, 0.0048)
    a = grid.cell_of((0.0, 0.0, 0.0))
    b = grid.cell_of((grid.cell, 0.0, 0.0))
    c = grid.cell_of((0.0, grid.cell, 0.0))
    assert b == a + 1 and c == a + grid.dims[0]


def test_grid_new_matches_the_reference_formula():  # sim.rs:24-32
    g = fa.Grid.new(HALF, 0.0048)
    for i in range(3):
        expect = int(math.ceil(np.float32(2.0) * np.float32(HALF[i]) / np.float32(0.0048))) + 2
        assert g.dims[i] == expect
        assert g.min[i] == float(np.float32(-np.float32(HALF[i]) - np.float32(0.0048)))


def test_cell_of_saturates_and_clamps():  # sim.rs:38-43 (Rust `as u32` saturates)
    g = fa.Grid.new((0.5, 0.5, 0.5), 0.1)
    assert g.cell_of((-10.0, -10.0, -10.0)) == 0
    far = g.cell_of((10.0, 10.0, 10.0))
    assert far == g.cell_count() - 1
    assert g.cell_of((float("nan"), 0.0, 0.0)) == g.cell_of((-10.0, 0.0, 0.0))


def test_tracers_seed_inside_the_fluid_region():  # sim.rs:485-494
    extent = (0.0357, 0.0774)
    half = (extent[0], extent[1], SLAB_DEPTH / 2)
    seeded = fa.seed_tracers(1000, half, 0.5)
    x = ((seeded[:, 0] & 0xFFFF) / 65535.0 * 2 - 1) * extent[0]
    y = ((seeded[:, 0] >> 16) / 65535.0 * 2 - 1) * extent[1]
    z = (seeded[:, 1] & 0xFFFF).astype(np.uint16).view(np.float16).astype(np.float32)
    assert np.all(np.abs(x) < extent[0]) and np.all(np.abs(z) < 0.5 * SLAB_DEPTH)
    assert np.all(y > -extent[1]) and np.all(y < 0.001)


def test_step_immediates_land_at_the_shader_offsets():  # sim.rs:496-546
    raw = fa.pack_step(force=(1, 2, 3), omega=(7, 8, 9), domega=(10, 11, 12), dt=4, v_clamp=5, seed=6,
                       touches=[(13, 14, 15, 16), (18, 19, 20, 21)], touch_n=2, touch_r=17)
    assert len(raw) == 144
    f = lambda o: float(np.frombuffer(raw, np.float32, 1, o)[0])  # noqa: E731
    u = lambda o: int(np.frombuffer(raw, np.uint32, 1, o)[0])  # noqa: E731
    assert [f(0), f(4), f(8)] == [1, 2, 3] and f(12) == 4
    assert [f(16), f(20), f(24)] == [7, 8, 9] and f(28) == 5
    assert [f(32), f(36), f(40)] == [10, 11, 12] and u(44) == 6
    assert f(48) == 17 and u(52) == 2
    assert [f(64), f(68), f(72), f(76)] == [13, 14, 15, 16]
    assert [f(80), f(84), f(88), f(92)] == [18, 19, 20, 21]


def test_sim_params_land_at_the_shader_offsets():  # sim.rs:548-567
    g = fa.Grid(min=(1.0, 2.0, 3.0), cell=4.0, dims=(5, 6, 7))
    raw = fa.pack_sim_params(g, 8, 9.0, 10.0, REST_DENSITY)
    f = lambda o: np.frombuffer(raw, np.float32, 1, o)[0]  # noqa: E731
    u = lambda o: int(np.frombuffer(raw, np.uint32, 1, o)[0])  # noqa: E731
    assert [f(0), f(4), f(8), f(12)] == [1, 2, 3, 4]
    assert [u(16), u(20), u(24), u(28)] == [5, 6, 7, 8]
    assert [f(32), f(36), f(40)] == [9, 10, np.float32(REST_DENSITY)]
    sigma = np.float32(1.0) / (np.float32(math.pi) * np.float32(729.0))
    assert [f(44), f(48), f(52), f(56)] == [np.float32(1) / np.float32(9), sigma, sigma / np.float32(9),
                                            np.float32(324)]
    assert raw[60:64] == b"\0\0\0\0"


def test_half_bits_round_to_nearest_even():  # sim.rs:135-151 vs IEEE binary16 (numpy)
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.uniform(-1, 1, 2000), rng.uniform(-100, 100, 2000),
                           rng.uniform(-1e-4, 1e-4, 2000), [0.0, 1.0, 0.5, 65000.0, 6.1e-5, 5.96e-8]])
    vals = vals.astype(np.float32)
    for v in vals:
        assert fa.half_bits(float(v)) == int(np.float16(v).view(np.uint16)), v


def test_pack_tracer_round_trips():  # sim.rs:124-175
    extent = (0.0357, 0.0774)
    for p in [(0.01, -0.02, 0.003), (-0.0357, 0.0774, -0.015), (0.0, 0.0, 0.0)]:
        raw = fa.pack_tracer(p, 0.25, extent)
        x = ((raw[0] & 0xFFFF) / 65535.0 * 2 - 1) * extent[0]
        y = ((raw[0] >> 16) / 65535.0 * 2 - 1) * extent[1]
        z = float(np.uint16(raw[1] & 0xFFFF).view(np.float16))
        s = float(np.uint16(raw[1] >> 16).view(np.float16))
        assert abs(x - p[0]) < extent[0] / 65535 * 1.01 and abs(y - p[1]) < extent[1] / 65535 * 1.01
        assert abs(z - p[2]) <= abs(p[2]) * 1e-3 + 1e-7 and abs(s - 0.25) < 1e-3


def v_for_cfl(count, dt, spacing):  # render.rs:3691-3693
    return float((np.float32(count) - np.float32(0.5)) * np.float32(0.4) * np.float32(spacing) / np.float32(dt))


NOMINAL = float(np.float32(1.0) / np.float32(120.0))


def test_the_substep_floor_scales_with_the_spacing():  # render.rs:3695-3702 via substeps_for(v=0)
    assert fa.substeps_for(NOMINAL, 0.0, 0.01, 16) == 2
    assert fa.substeps_for(NOMINAL, 0.0, 0.0062, 16) == 4
    assert fa.substeps_for(0.033, 0.0, 0.01, 16) == 8
    assert fa.substeps_for(0.033, 0.0, 0.0062, 16) == 8


def test_a_120hz_frame_jumps_six_or_seven_substeps_to_eight():  # render.rs:3704-3711
    counts = [fa.substeps_for(NOMINAL, v_for_cfl(k, NOMINAL, 0.01), 0.01, 16) for k in range(5, 10)]
    assert counts == [5, 8, 8, 8, 9]


def test_a_120hz_frame_at_4x_floors_at_four_then_jumps():  # render.rs:3713-3718
    assert fa.substeps_for(NOMINAL, v_for_cfl(1, NOMINAL, 0.0062), 0.0062, 16) == 4
    assert fa.substeps_for(NOMINAL, v_for_cfl(6, NOMINAL, 0.0062), 0.0062, 16) == 8


def test_a_slipped_frame_keeps_its_cfl_count_and_a_flung_one_stops_at_the_cap():  # render.rs:3720-3729
    assert fa.substeps_for(0.010, v_for_cfl(7, 0.010, 0.01), 0.01, 16) == 7
    assert fa.substeps_for(NOMINAL, 1.0e6, 0.01, 16) == 16


def test_refine_passes():  # render.rs:77-79
    assert fa.refine_passes(0.00106) == 5 and fa.refine_passes(0.00105) == 2


@pytest.mark.parametrize("g,a,expect", [((0, 0, -1), (0, 0, 0), (0.0, 0.0, -9.80665)),
                                        ((0, 0, -1), (-0.5, 0, 0), None),
                                        ((0, 0, -1), (0, 0, 1), (0.0, 0.0, 0.0))])
def test_body_force(g, a, expect):  # lib.rs:43-72
    f = fa.body_force(g, a)
    if expect is None:
        assert f[0] == float(np.float32(-0.5) * np.float32(9.80665))
    else:
        assert f == tuple(float(np.float32(x)) for x in expect)


def test_spacing_for_lands_one_x_on_the_shipped_spacing():  # render.rs:1355-1380
    extent = (4.0 * 0.0357, 4.0 * 0.0774)
    assert abs(fa.spacing_for(1.0, extent) - 0.01) < 1e-7
    assert fa.spacing_for(4.0, extent) < fa.spacing_for(1.0, extent)
