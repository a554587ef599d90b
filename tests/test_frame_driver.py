"""The frame driver's input filters and lens ends (include/fluid_sim.h,
fluid_filters_*, fluid_lens_ends), mirroring the reference's own unit tests
of them (crates/fluid-core/src/render.rs:3582-3687) and the definitions they
pin (render.rs:649-680, 862-1105). Pure host code: no device needed.
"""
import ctypes as C
import math

import numpy as np
import pytest

import fluid_amd as fa


def f3(*v):
    return (C.c_float * 3)(*[float(x) for x in v])


class Filters:
    def __init__(self):
        self.L = fa.lib()
        self.h = self.L.fluid_filters_create()

    def __del__(self):
        self.L.fluid_filters_destroy(self.h)

    def asleep(self, touched, smooth, dev, omega, v_max):
        return bool(self.L.fluid_filters_asleep(self.h, touched, f3(*smooth), dev, f3(*omega), v_max))

    def force(self, body):
        out = (C.c_float * 3)()
        dev = self.L.fluid_filters_force(self.h, f3(*body), out)
        return list(out), dev

    def rotation(self, rate, dt):
        o, d = (C.c_float * 3)(), (C.c_float * 3)()
        self.L.fluid_filters_rotation(self.h, f3(*rate), dt, o, d)
        return list(o), list(d)

    def touch(self, slot, at):
        if at is None:
            self.L.fluid_filters_touch(self.h, slot, 0.0, 0.0, False)
        else:
            self.L.fluid_filters_touch(self.h, slot, at[0], at[1], True)

    def fingers(self, extent, dt):
        t = (C.c_float * 20)()
        r = C.c_float()
        n = self.L.fluid_filters_fingers(self.h, (C.c_float * 2)(*extent), dt, t, C.byref(r))
        return n, np.array(t, np.float32).reshape(5, 4), r.value


STILL = None


def still_frames():
    return int(fa.lib().fluid_filters_still_frames())


def test_idle_gate_sleeps_when_still_and_wakes_on_shake_or_tilt():  # render.rs:3582-3603
    g = 9.81
    still = (0.0, 0.0, 0.0)
    upright = (0.0, -g, 0.0)
    gate = Filters()
    for _ in range(still_frames() - 1):
        assert not gate.asleep(False, upright, 0.1, still, 0.03)
    assert gate.asleep(False, upright, 0.1, still, 0.03)
    # noise peaks between the sleep and wake thresholds: no false wake
    assert gate.asleep(False, upright, 0.8, still, 0.03)
    # a shake crosses the deviation test in one tick
    assert not gate.asleep(False, upright, 2.0, still, 0.03)
    for _ in range(still_frames()):
        gate.asleep(False, upright, 0.1, still, 0.03)
    assert gate.asleep(False, upright, 0.1, still, 0.03)
    # a settled 2.9-degree tilt is past the 1.5-degree wake angle
    assert not gate.asleep(False, (g * 0.05, -g, 0.0), 0.1, still, 0.03)


def test_idle_gate_wakes_on_pure_rotation():  # render.rs:3608-3629
    flat = (0.0, 0.0, -9.81)
    gate = Filters()
    for _ in range(still_frames()):
        gate.asleep(False, flat, 0.1, (0.0, 0.0, 0.0), 0.03)
    assert gate.asleep(False, flat, 0.1, (0.0, 0.0, 0.01), 0.03), "gyro noise"
    assert not gate.asleep(False, flat, 0.1, (0.0, 0.0, 0.5), 0.03), "a real turn"
    spun = Filters()
    for _ in range(2 * still_frames()):
        assert not spun.asleep(False, flat, 0.1, (0.0, 0.0, 0.5), 0.03)


def test_idle_gate_wakes_on_a_touch_and_on_moving_water():
    """A finger down wakes the gate at once; water above the sleep speed
    (0.04 m/s, render.rs:1030) keeps it from counting still frames."""
    gate = Filters()
    up = (0.0, -9.81, 0.0)
    for _ in range(still_frames()):
        gate.asleep(False, up, 0.1, (0.0, 0.0, 0.0), 0.03)
    assert gate.asleep(False, up, 0.1, (0.0, 0.0, 0.0), 0.03)
    assert not gate.asleep(True, up, 0.1, (0.0, 0.0, 0.0), 0.03)
    busy = Filters()
    for _ in range(2 * still_frames()):
        assert not busy.asleep(False, up, 0.1, (0.0, 0.0, 0.0), 0.05)


def test_the_fingers_map_drawable_points_into_the_box():  # render.rs:3633-3652
    extent = (0.1, 0.2)
    f = Filters()
    f.touch(0, (1.0, 0.0))
    n, t, radius = f.fingers(extent, 0.0)
    assert n == 1
    assert tuple(t[0, :2]) == (np.float32(0.1), np.float32(0.2)), "the drawable's top right is the box's"
    assert tuple(t[0, 2:]) == (0.0, 0.0), "one sample makes no speed"
    assert radius > 0.0
    f.touch(0, None)
    assert f.fingers(extent, 1.0 / 120.0)[0] == 0, "lifted"


def test_a_lifted_finger_leaves_the_others_where_they_were():  # render.rs:3654-3687
    extent = (0.1, 0.2)
    dt = 1.0 / 120.0
    f = Filters()
    velocity = None
    for fr in range(60):
        f.touch(0, (0.9, 0.9))
        f.touch(2, (0.2 + 0.01 * fr, 0.5))
        n, t, _ = f.fingers(extent, dt)
        assert n == 2
        velocity = t[1, 2:]
    assert abs(velocity[0] - 0.24) < 0.005, velocity
    assert velocity[1] == 0.0
    f.touch(0, None)
    f.touch(2, (0.2 + 0.01 * 60.0, 0.5))
    n, t, _ = f.fingers(extent, dt)
    assert n == 1
    assert abs(t[0, 2] - 0.24) < 0.005, t[0]


def test_force_filter_follows_a_step_at_its_deviation_rate():
    """ForceFilter (render.rs:878-903): the first sample primes it; each
    later one moves the smoothed force by alpha = clamp(|dev| / 2, 0.02, 1)
    of the gap, |dev| the gap before the move."""
    f = Filters()
    out, dev = f.force((0.0, -9.81, 0.0))
    assert out == [0.0, np.float32(-9.81), 0.0] and dev == 0.0
    out, dev = f.force((1.0, -9.81, 0.0))  # a 1 m/s^2 gap: alpha 0.5
    assert abs(dev - 1.0) < 1e-6 and abs(out[0] - 0.5) < 1e-6
    out, dev = f.force((1.0, -9.81, 0.0))  # gap 0.5: alpha 0.25
    assert abs(dev - 0.5) < 1e-6 and abs(out[0] - 0.625) < 1e-6
    gap = 1.0 - out[0]
    for _ in range(200):
        out, dev = f.force((1.0, -9.81, 0.0))
        # past a 0.04 gap alpha sits on its 0.02 floor: the gap shrinks by 2% a frame
        assert 0.0 < 1.0 - out[0] <= gap * (1.0 - 0.02) + 1e-7
        gap = 1.0 - out[0]
    assert gap < 2e-3
    out, dev = f.force((5.0, -9.81, 0.0))  # a shake: alpha 1
    assert abs(out[0] - 5.0) < 1e-6


def test_rotation_tracker_smooths_and_differentiates():
    """RotationTracker (render.rs:1001-1015): the first sample, or a gap
    outside (0, 0.1] s, primes it with zero derivative; otherwise the rate
    chases with alpha = 1 - exp(-dt / 25 ms) and the derivative is the
    smoothed change over dt, clamped to +-200 rad/s^2."""
    f = Filters()
    o, d = f.rotation((1.0, 0.0, 0.0), 1.0 / 120.0)
    assert o == [1.0, 0.0, 0.0] and d == [0.0, 0.0, 0.0]
    dt = 1.0 / 120.0
    a = 1.0 - math.exp(-dt / 0.025)
    o, d = f.rotation((2.0, 0.0, 0.0), dt)
    assert abs(o[0] - (1.0 + a)) < 1e-6
    assert abs(d[0] - a / dt) < 1e-3
    o, d = f.rotation((200.0, 0.0, 0.0), dt)
    assert d[0] == 200.0  # clamped
    o, d = f.rotation((0.0, 0.0, 0.0), 0.5)  # a long gap re-primes
    assert o == [0.0, 0.0, 0.0] and d == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("lens,idx", [(0, (11, 6)), (1, (12, 13)), (2, (4, 5)), (3, (2, 3)), (4, (11, 6)),
                                      (9, (11, 6))])
def test_lens_ends_come_from_the_stats_block(lens, idx):
    """Lens::ends (render.rs:649-656): velocity and direction read the speed
    pair, acceleration and pressure their lens EMAs' pairs, proximity the
    density pair; an unknown code paints velocity (lib.rs:215-221)."""
    stats = (C.c_float * 14)(*[float(10 * i + 1) for i in range(14)])
    ends, floor = (C.c_float * 2)(), C.c_float()
    assert fa.lib().fluid_lens_ends(lens, stats, 0.01, ends, C.byref(floor)) == 0
    assert (ends[0], ends[1]) == (stats[idx[0]], stats[idx[1]])


def test_lens_floors_scale_with_the_spacing():
    """Lens::floor (render.rs:658-680): sqrt(2 g d), g, rho0 g d, 1% of rho0."""
    g, rho0, d = 9.80665, 998.2, 0.01
    want = {0: math.sqrt(2 * g * d), 1: g, 2: rho0 * g * d, 3: 0.01 * rho0, 4: math.sqrt(2 * g * d)}
    stats = (C.c_float * 14)()
    for lens, w in want.items():
        floor = C.c_float()
        fa.lib().fluid_lens_ends(lens, stats, d, None, C.byref(floor))
        assert abs(floor.value - w) <= 1e-6 * w, lens
