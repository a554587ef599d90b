"""The renderer ABI's frame driver on the GPU against the oracle and the
reference's rules:

- frames of fluid_renderer_frame (filters, CFL schedule, fingers, stats,
  tracers) equal the oracle replaying the Step blocks the driver used
  (fluid_renderer_last_step), word for word, tracers included;
- fluid_sim_device_records hands out the resting set the sprite and disc
  pullers bind (render.rs:1818-1828): the same words as a download;
- the stage-0 microbench (bench_sweeps > 0, render.rs:2429-2729) validates
  its first frame as validate_bench does, its starts equal numpy's counting
  scan and its densities a float64 brute force to 1e-5;
- a flat look's ramp follows Lens::ends of the stats block (render.rs:
  649-743).
"""
import ctypes as C
import math

import numpy as np
import pytest

import fluid_amd as fa
from fluid_amd._lib import check, lib
from oracle import OracleSim

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")
FRAME_MS = 1000.0 / 60.0


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class DeviceRecords(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("positions", "velocities", "prev_vel", "prev_pressure", "temperature",
                                          "density", "tracers")] + \
               [("count", C.c_uint32), ("tracer_count", C.c_uint32), ("stream", C.c_void_p)]


def _abi():
    L = lib()
    vp = C.c_void_p
    L.fluid_renderer_create.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_float,
                                        C.c_uint32, C.c_uint32, C.c_float]
    L.fluid_renderer_create.restype = vp
    L.fluid_renderer_frame.argtypes = [vp, Vec3, Vec3, Vec3, C.c_double]
    L.fluid_renderer_frame.restype = C.c_uint32
    L.fluid_renderer_touch.argtypes = [vp, C.c_uint32, C.c_float, C.c_float, C.c_bool]
    L.fluid_renderer_touch.restype = None
    L.fluid_renderer_set_look.argtypes = [vp, C.c_bool, C.c_bool, Vec3, C.c_bool, Vec3, C.c_uint32]
    L.fluid_renderer_set_look.restype = None
    L.fluid_renderer_destroy.argtypes = [vp]
    L.fluid_renderer_destroy.restype = None
    return L


def _sim_state(L, sim, n):
    out = {k: np.empty((n, 4) if k in ("pos", "vel", "prev_vel") else n, np.float32) for k in STATE}
    f = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    check(L.fluid_sim_download_state(sim, *[f(out[k]) for k in STATE]), "download_state")
    return out


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _frame(L, r, k, touch_from=None):
    th = 0.4 * math.sin(0.07 * k)
    zero = Vec3(0.0, 0.0, 0.0)
    if touch_from is not None and k >= touch_from:
        L.fluid_renderer_touch(r, 0, 0.3 + 0.02 * (k - touch_from), 0.85, True)
    return L.fluid_renderer_frame(r, Vec3(math.sin(th), -math.cos(th), 0.0), zero, Vec3(0.0, 0.0, 0.3 * th),
                                  k * FRAME_MS)


def test_renderer_frames_equal_the_oracle_on_the_same_steps():
    L = _abi()
    r = L.fluid_renderer_create(None, 1170, 2532, 0, 0.0, 0, 0.0, 8, 256, 1.0)
    assert r
    try:
        sim = L.fluid_renderer_sim(r)
        n = int(L.fluid_sim_count(sim))
        ntr = int(L.fluid_sim_tracer_count(sim))
        params = (C.c_uint8 * 64)()
        check(L.fluid_sim_params(sim, params), "params")
        print(f"renderer: {n} particles, {ntr} tracers, 40 frames at 60 Hz, a finger from frame 12", flush=True)
        start = _sim_state(L, sim, n)
        tr = np.empty((ntr, 2), np.uint32)
        check(L.fluid_sim_download_tracers(sim, tr.ctypes.data_as(C.POINTER(C.c_uint32))), "tracers")
        o = OracleSim(bytes(params))
        o.set_state(*[start[k] for k in STATE])
        step, tstep = (C.c_uint8 * 144)(), (C.c_uint8 * 144)()
        nsub, refine = C.c_uint32(), C.c_uint32()
        stepped = 0
        for k in range(40):
            _frame(L, r, k, touch_from=12)
            check(L.fluid_renderer_last_step(r, step, C.byref(nsub), C.byref(refine), tstep), "last_step")
            if nsub.value == 0:
                continue
            for _ in range(nsub.value):
                o.substep(bytes(step), refine.value)
            tr = o.tracers(bytes(tstep), tr)
            stepped += nsub.value
        got = _sim_state(L, sim, n)
        gtr = np.empty((ntr, 2), np.uint32)
        check(L.fluid_sim_download_tracers(sim, gtr.ctypes.data_as(C.POINTER(C.c_uint32))), "tracers")
        print(f"replayed {stepped} substeps", flush=True)
        assert stepped > 40
        ost = o.state()
        for key in STATE:
            assert np.array_equal(_bits(got[key]), _bits(ost[key])), key
        assert np.array_equal(gtr, tr)
        assert not np.array_equal(got["pos"], start["pos"])
    finally:
        L.fluid_renderer_destroy(r)


def _hip():
    lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.restype = C.c_int
    return h


def test_device_records_are_the_resting_set():
    sc = fa.scenes.small_block((10, 10, 10), seed=3, tracers=300)
    n, dt, refine, vc = sc.schedule()
    h = _hip()
    with fa.Sim(sc.params, tracers=len(sc.tracers)) as s:
        s.upload_state(sc.pos4)
        s.upload_tracers(sc.tracers)
        s.frame(sc.step_block(0.0, dt, vc), 3, refine, fa.pack_step(force=(0, -9.81, 0), dt=3 * dt, seed=2))
        d = DeviceRecords()
        check(s.L.fluid_sim_device_records(s.h, C.byref(d)), "device_records")
        assert d.count == sc.n and d.tracer_count == len(sc.tracers)
        assert h.hipStreamSynchronize(d.stream) == 0
        want = s.state()
        for key, ptr, shape in (("pos", d.positions, (sc.n, 4)), ("vel", d.velocities, (sc.n, 4)),
                                ("prev_vel", d.prev_vel, (sc.n, 4)), ("prev_pressure", d.prev_pressure, (sc.n,)),
                                ("temperature", d.temperature, (sc.n,))):
            a = np.empty(shape, np.float32)
            assert h.hipMemcpy(a.ctypes.data, ptr, a.nbytes, 2) == 0
            assert np.array_equal(_bits(a), _bits(want[key])), key
        a = np.empty(sc.n, np.float32)
        assert h.hipMemcpy(a.ctypes.data, d.density, a.nbytes, 2) == 0
        assert np.array_equal(_bits(a), _bits(s.field("density")))
        t = np.empty((len(sc.tracers), 2), np.uint32)
        assert h.hipMemcpy(t.ctypes.data, d.tracers, t.nbytes, 2) == 0
        assert np.array_equal(t, s.download_tracers())


@pytest.mark.parametrize("sweeps,spacing", [(1, 0.0), (3, 0.012)])
def test_stage0_bench_validates_and_matches_brute_force(sweeps, spacing):
    L = _abi()
    r = L.fluid_renderer_create(None, 1170, 2532, 0, 0.0, sweeps, spacing, 0, 0, 1.0)
    assert r, "the stage-0 bench mode returned NULL"
    try:
        cnt, cells = C.c_uint32(), C.c_uint32()
        assert L.fluid_renderer_bench(r, None, None, C.byref(cnt), C.byref(cells), None, None, None) == 0
        n, c = cnt.value, cells.value
        print(f"stage-0 bench: {n} particles, {c} cells, {sweeps} sweeps", flush=True)
        zero = Vec3(0.0, 0.0, 0.0)
        assert L.fluid_renderer_frame(r, Vec3(0.0, -1.0, 0.0), zero, zero, 0.0) == 1
        ok_scan, ok_dens = C.c_int(), C.c_int()
        rho = np.empty(n, np.float32)
        starts = np.empty(c + 1, np.uint32)
        pos = np.empty((n, 4), np.float32)
        assert L.fluid_renderer_bench(r, C.byref(ok_scan), C.byref(ok_dens), None, None,
                                      rho.ctypes.data_as(C.POINTER(C.c_float)),
                                      starts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      pos.ctypes.data_as(C.POINTER(C.c_float))) == 0
        assert (ok_scan.value, ok_dens.value) == (1, 1)
        # the counting scan of the seeded positions in the bench's grid
        # (render.rs:2461-2464: h = 1.2 d, support 2h over the pixel extent)
        f = np.float32
        mpp = f(f(4.0) * f(0.0254)) / f(458.0)
        d = f(spacing) if spacing > 0 else f(0.8) * f(f(4.0) * f(0.0025))
        h = f(1.2) * d
        half = (float(f(1170.0) * f(0.5) * mpp), float(f(2532.0) * f(0.5) * mpp), float(f(lib().fluid_slab_depth()) * f(0.5)))
        grid = fa.Grid.new(half, float(f(2.0) * h))
        assert grid.cell_count() == c
        p = pos[:, :3]
        mn, cell, dims = np.array(grid.min, np.float32), np.float32(grid.cell), np.array(grid.dims)
        qc = (p - mn) / cell
        coord = np.minimum(np.where(qc > 0, np.floor(qc), 0).astype(np.int64), dims - 1)
        cid = (coord[:, 2] * dims[1] + coord[:, 1]) * dims[0] + coord[:, 0]
        assert np.array_equal(starts, np.concatenate([[0], np.cumsum(np.bincount(cid, minlength=c))]))
        # the density of every particle against a float64 brute force of the
        # cubic spline (sim.rs:49-60), mass rho0 d^3
        h, m = float(h), 998.2 * float(d) ** 3
        x = p.astype(np.float64)
        rr = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
        q = rr / h
        w = np.where(q < 1, 1 - 1.5 * q * q * (1 - 0.5 * q), np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0)) / (math.pi * h ** 3)
        want = m * w.sum(1)
        err = np.abs(rho - want) / want
        assert err.max() < 1e-5, err.max()
        assert starts[-1] == n and np.all(np.diff(starts.astype(np.int64)) >= 0)
        # a second frame sorts and sweeps again: the same bits
        assert L.fluid_renderer_frame(r, Vec3(0.0, -1.0, 0.0), zero, zero, FRAME_MS) == 1
        again = np.empty(n, np.float32)
        L.fluid_renderer_bench(r, None, None, None, None, again.ctypes.data_as(C.POINTER(C.c_float)), None, None)
        assert np.array_equal(_bits(again), _bits(rho))
        assert L.fluid_renderer_sim(r) is None
    finally:
        L.fluid_renderer_destroy(r)


def test_a_flat_look_ramp_follows_the_lens_ends():
    """Lens 0 (velocity) on the flat look: no ramp before a stats block with
    density has landed; the first live ramp is Lens::ends of the last
    readback with the high end lifted to low + floor (render.rs:716-726);
    later ends chase and keep the floor; a change of lens restarts on the
    new lens's ends."""
    L = _abi()
    r = L.fluid_renderer_create(None, 1170, 2532, 0, 0.0, 0, 0.0, 8, 0, 1.0)
    assert r
    try:
        L.fluid_renderer_set_look(r, True, False, Vec3(0.1, 0.2, 0.9), True, Vec3(0.9, 0.9, 1.0), 0)
        sim = L.fluid_renderer_sim(r)
        ends = (C.c_float * 2)()
        spacing = None
        stats_prev = None
        for k in range(12):
            _frame(L, r, k)
            live = L.fluid_renderer_ramp(r, ends)
            if k <= 1:
                assert live == 0, k  # frame 0 steps nothing; frame 1 follows zero stats
            else:
                assert live == 1, k
                want, floor = (C.c_float * 2)(), C.c_float()
                if spacing is None:
                    # the spacing of the renderer's scene: d = h / 1.2, h at bytes 32-35 of SimParams
                    params = (C.c_uint8 * 64)()
                    L.fluid_sim_params(sim, params)
                    spacing = float(np.frombuffer(bytes(params), np.float32, 1, 32)[0]) / 1.2
                L.fluid_lens_ends(0, stats_prev, spacing, want, C.byref(floor))
                assert ends[1] >= ends[0] + floor.value * (1 - 1e-6)
                if k == 2:  # the first live ramp takes the ends whole
                    assert ends[0] == want[0]
                    assert abs(ends[1] - max(want[1], want[0] + floor.value)) <= 1e-6 * abs(ends[1])
            stats_prev = (C.c_float * 14)()
            L.fluid_sim_read_stats(sim, stats_prev)
        L.fluid_renderer_set_look(r, True, True, Vec3(0.1, 0.2, 0.9), True, Vec3(0.9, 0.9, 1.0), 3)
        _frame(L, r, 12)
        assert L.fluid_renderer_ramp(r, ends) == 1
        want = (C.c_float * 2)()
        L.fluid_lens_ends(3, stats_prev, spacing, want, None)
        assert ends[0] == want[0]  # proximity: the density pair, taken whole on the switch
        L.fluid_renderer_set_look(r, True, False, Vec3(0.1, 0.2, 0.9), False, Vec3(0.9, 0.9, 1.0), 0)
        _frame(L, r, 13)
        assert L.fluid_renderer_ramp(r, ends) == 0  # a solid paint has no ramp
    finally:
        L.fluid_renderer_destroy(r)
