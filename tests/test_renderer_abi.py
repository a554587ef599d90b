"""The reference's renderer ABI (include/fluid_ffi.h), driven the way the
app's FrameDriver drives it, through ctypes: create, frames, touches,
reseeding, stats and destroy, headless on the HIP step.

The frame driver's host pieces (dt clamp, substep count, refine passes,
body force) are pinned against the reference's rules in
test_host_helpers.py; here the ABI must step, report sane stats, stay
deterministic, and refuse the mode it does not serve (the sprite demo). The
refusal and the NULL-handle paths touch no device and run without a GPU.
"""
import ctypes as C
import math

import pytest

from fluid_amd._lib import lib

class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("frames", C.c_uint64)] + [(n, C.c_float) for n in (
        "interval_p50_us", "interval_p99_us", "interval_max_us", "acquire_p50_us", "acquire_p99_us",
        "encode_p50_us", "encode_p99_us", "gpu_p50_us", "gpu_p99_us", "compression_avg", "compression_max",
        "density_min", "density_max", "pressure_min", "pressure_max", "v_max", "temperature_min",
        "temperature_max")] + [("clamp_count", C.c_uint32), ("neighbour_overflow", C.c_uint32),
                               ("substeps", C.c_uint32), ("idle_frames", C.c_uint64)]


PHYS = ("density_min", "density_max", "pressure_min", "pressure_max", "v_max", "temperature_min",
        "temperature_max", "compression_avg", "compression_max", "clamp_count", "neighbour_overflow", "substeps")
FRAME_MS = 1000.0 / 60.0


def _abi():
    L = lib()
    vp = C.c_void_p
    L.fluid_renderer_create.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_float,
                                        C.c_uint32, C.c_uint32, C.c_float]
    L.fluid_renderer_create.restype = vp
    L.fluid_renderer_frame.argtypes = [vp, Vec3, Vec3, Vec3, C.c_double]
    L.fluid_renderer_frame.restype = C.c_uint32
    L.fluid_renderer_set_particles.argtypes = [vp, C.c_float]
    L.fluid_renderer_set_particles.restype = None
    L.fluid_renderer_particles_at.argtypes = [vp, C.c_float]
    L.fluid_renderer_particles_at.restype = C.c_uint32
    L.fluid_renderer_touch.argtypes = [vp, C.c_uint32, C.c_float, C.c_float, C.c_bool]
    L.fluid_renderer_touch.restype = None
    L.fluid_renderer_set_look.argtypes = [vp, C.c_bool, C.c_bool, Vec3, C.c_bool, Vec3, C.c_uint32]
    L.fluid_renderer_set_look.restype = None
    L.fluid_renderer_resize.argtypes = [vp, C.c_uint32, C.c_uint32]
    L.fluid_renderer_resize.restype = None
    L.fluid_renderer_stats.argtypes = [vp]
    L.fluid_renderer_stats.restype = Stats
    L.fluid_renderer_destroy.argtypes = [vp]
    L.fluid_renderer_destroy.restype = None
    return L


def _create(L, scale=1.0, substeps=8, tracers=64, spacing=0.0):
    r = L.fluid_renderer_create(None, 1170, 2532, 0, 0.0, 0, spacing, substeps, tracers, scale)
    assert r, "fluid_renderer_create returned NULL"
    return r


def _run(L, r, frames, start=0, touch_from=None):
    """`frames` frames at 60 Hz under a slowly tilting gravity (in g); with
    `touch_from`, finger 0 goes down at that frame and drags right."""
    zero = Vec3(0.0, 0.0, 0.0)
    stepped = []
    for k in range(start, start + frames):
        th = 0.3 * math.sin(0.05 * k)
        if touch_from is not None and k >= touch_from:
            L.fluid_renderer_touch(r, 0, 0.3 + 0.02 * (k - touch_from), 0.8, True)
        stepped.append(L.fluid_renderer_frame(r, Vec3(math.sin(th), -math.cos(th), 0.0), zero, zero, k * FRAME_MS))
    return stepped


def _phys(s):
    return {k: getattr(s, k) for k in PHYS}


def test_the_sprite_demo_mode_returns_null():
    """The M2 sprite demo (sim_substeps = 0, no bench sweeps) is not served;
    the stage-0 microbench is (tests/test_gpu_frame_driver.py)."""
    L = _abi()
    assert not L.fluid_renderer_create(None, 1170, 2532, 4096, 3.0, 0, 0.0, 0, 0, 1.0)  # sprite demo


def test_null_handle_is_harmless():
    L = _abi()
    zero = Vec3(0.0, 0.0, 0.0)
    assert L.fluid_renderer_frame(None, zero, zero, zero, 0.0) == 0
    assert L.fluid_renderer_stats(None).frames == 0
    assert L.fluid_renderer_particles_at(None, 1.0) == 0
    L.fluid_renderer_touch(None, 0, 0.5, 0.5, True)
    L.fluid_renderer_destroy(None)


def test_touch_slots_is_the_reference_constant():
    assert C.c_uint32.in_dll(lib(), "FLUID_TOUCH_SLOTS").value == 5


@pytest.mark.gpu
def test_frames_step_and_report_sane_stats():
    L = _abi()
    r = _create(L)
    try:
        assert L.fluid_renderer_frame(r, Vec3(0, -1, 0), Vec3(0, 0, 0), Vec3(0, 0, 0), 0.0) == 1
        s = L.fluid_renderer_stats(r)
        assert s.frames == 1 and s.substeps == 0  # the first frame has no interval: dt = 0
        assert all(x == 1 for x in _run(L, r, 30, start=1))
        s = L.fluid_renderer_stats(r)
        assert s.frames == 31
        assert 1 <= s.substeps <= 8
        assert 0.0 < s.density_min <= s.density_max < 5000.0
        assert s.pressure_min >= 0.0 and math.isfinite(s.pressure_max)
        assert 0.0 < s.v_max < 10.0
        assert s.neighbour_overflow == 0
        assert s.gpu_p50_us > 0.0 and s.interval_p50_us == pytest.approx(FRAME_MS * 1000.0, rel=1e-3)
        assert s.acquire_p50_us == 0.0 and s.encode_p50_us == 0.0  # headless: nothing acquired or encoded
    finally:
        L.fluid_renderer_destroy(r)


@pytest.mark.gpu
def test_same_inputs_give_the_same_stats():
    L = _abi()
    out = []
    for _ in range(2):
        r = _create(L)
        _run(L, r, 20)
        out.append(_phys(L.fluid_renderer_stats(r)))
        L.fluid_renderer_destroy(r)
    assert out[0] == out[1]


@pytest.mark.gpu
def test_a_dragged_finger_moves_the_fluid():
    L = _abi()
    runs = []
    for touch in (None, 5):
        r = _create(L)
        _run(L, r, 20, touch_from=touch)
        runs.append(L.fluid_renderer_stats(r))
        L.fluid_renderer_destroy(r)
    assert runs[1].v_max != runs[0].v_max
    assert runs[1].v_max > 0.0


@pytest.mark.gpu
def test_reseeding_changes_the_count_and_keeps_stepping():
    L = _abi()
    r = _create(L, scale=0.25)
    try:
        small, big = L.fluid_renderer_particles_at(r, 0.25), L.fluid_renderer_particles_at(r, 1.0)
        assert 0 < small < big
        _run(L, r, 5)
        L.fluid_renderer_set_particles(r, 1.0)
        L.fluid_renderer_resize(r, 1170, 2532)
        L.fluid_renderer_set_look(r, True, False, Vec3(0, 0, 0), False, Vec3(1, 1, 1), 0)
        assert all(x == 1 for x in _run(L, r, 5, start=5))
        s = L.fluid_renderer_stats(r)
        assert s.frames == 10 and s.density_min > 0.0 and s.neighbour_overflow == 0
    finally:
        L.fluid_renderer_destroy(r)
