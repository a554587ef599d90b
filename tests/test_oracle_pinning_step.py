"""Pins the oracle's remaining step pieces against what the mathematics fixes.

The density sweep, the grid and the wall polynomials are pinned in
test_oracle_pinning.py. This file pins the rest of the step, each piece by
something other than a re-reading of the reference shader:

  * integrate (sim_solve.wgsl:709-756): a lone particle in free flight
    follows the semi-implicit Euler closed form; the speed clamp lands on
    its ceiling (3 v_clamp below half rest density, v_clamp above); the wall
    clamp zeroes the normal component only; the acceleration lens is the
    exponential moving average of |g| in closed form; the temperature's
    pressure work is the isentropic heating dT = T beta / (rho c_p) dp of
    water's published constants;
  * forces_den_apply (sim_solve.wgsl:103-127, 630-663): the rotating frame's
    Euler, centrifugal and Coriolis terms from the vector identities; a
    finger's blend 1 - exp(-40 w dt) toward the weighted finger velocity;
  * forces_eval (sim_solve.wgsl:524-592): an O(N^2) float64 brute force of
    Morris et al. (1997) viscosity, Monaghan (1989) XSPH, Clavet et al.
    (2005) near pressure, Akinci et al. (2013) cohesion and wall adhesion,
    Cleary & Monaghan (1999) conduction and viscous heating; conduction
    conserves heat, viscosity dissipates kinetic energy;
  * div_apply / den_apply and den_kappa (sim_solve.wgsl:596-621, 674-698):
    the DFSPH updates of Bender & Koschier (2015) in O(N^2) float64, and a
    compressed block's density error falling pass by pass;
  * reduce_stats (sim_solve.wgsl:800-874): numpy's min, max and mean of the
    fields it reduces.

Constants are the reference's f32 values (the WGSL `const`s, sim_solve.wgsl:
84-172), as float64 numbers.
"""
import math

import numpy as np
import pytest

import fluid_amd as fa
from oracle import OracleSim, lib as olib

F32 = lambda x: float(np.float32(x))  # noqa: E731  a WGSL f32 constant as a float64
MU = F32(1.002e-3)
K_NEAR = 3000.0
HEAT_CAPACITY = F32(4184.0)
CONDUCTIVITY = F32(0.598)
EXPANSION = F32(2.07e-4)
SIGMA_WATER = F32(0.0728)
LATTICE_SIGMA = F32(0.8665)
COS_CONTACT = F32(-0.342)
PI_REF = 3.14159265  # the cohesion spline's literal (sim_solve.wgsl:536)


def _p(sc, off):
    return float(np.frombuffer(sc.params, np.float32, 1, off)[0])


def _consts(sc):
    return dict(h=_p(sc, 32), m=_p(sc, 36), rho0=_p(sc, 40), inv_h=_p(sc, 44), sig=_p(sc, 48), sig_h=_p(sc, 52),
                sup2=_p(sc, 56), cell=_p(sc, 12), bmin=np.frombuffer(sc.params, np.float32, 3, 0).astype(np.float64))


def w_shape(q):
    """The cubic spline's shape in q = r/h (W = sigma * w_shape, sigma = 1/(pi h^3))."""
    q = np.asarray(q, np.float64)
    return np.where(q < 1, 1 - 1.5 * q * q * (1 - 0.5 * q), np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


def dw_shape(q):
    q = np.asarray(q, np.float64)
    return np.where(q < 1, -3 * q + 2.25 * q * q, np.where(q < 2, -0.75 * (2 - q) ** 2, 0.0))


def _pairs(C, pos):
    """Pair geometry of all particles: x_ij, r, W_ij, grad_i W_ij, support mask."""
    x = pos[:, None, :] - pos[None, :, :]
    r2 = (x ** 2).sum(-1)
    inside = (r2 < C["sup2"]) & ~np.eye(len(pos), dtype=bool)
    r = np.sqrt(r2)
    q = r * C["inv_h"]  # q = r / h with the SimParams block's f32 1/h (sim.rs:343, sim_solve.wgsl:448)
    W = np.where(inside, C["sig"] * w_shape(q), 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        gradW = np.where(inside[..., None], (C["sig_h"] * dw_shape(q) / r)[..., None] * x, 0.0)
    return x, r, q, W, gradW, inside


def _lone(pos=(0.03, 0.01, -0.02), half=(0.2, 0.2, 0.2)):
    sc = fa.scenes.small_block((1, 1, 1), seed=0, spacing=0.01, half=half)
    sc.pos4[0, :3] = np.array(pos, np.float32)
    return sc


def _step(**kw):
    return fa.pack_step(**kw)


def _run(sc, precision, kernels, step, vel=None, prev_vel=None, prev_pressure=None, temperature=None, phys=None,
         cap=96, faithful=False):
    o = OracleSim(sc.params, nbr_cap=cap, precision=precision, faithful=faithful)
    if phys is not None:
        o.set_physics(*phys)
    o.set_state(sc.pos4, vel, prev_vel, prev_pressure, temperature)
    for k in kernels:
        o.kernel(k, step)
    return o


SUBSTEP = ("grid", "density_div", "div_apply", "forces_eval", "forces_den_apply") + \
    ("den_kappa", "den_apply") * 5 + ("integrate",)


# ---- integrate ----------------------------------------------------------------

@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 2e-6)])
def test_free_flight_is_semi_implicit_euler(precision, tol):
    """A lone particle far from the walls feels only the body force: after n
    substeps v_n = v0 + n dt g and x_n = x0 + n dt v0 + dt^2 g n (n + 1) / 2
    (velocity first, then position, sim_solve.wgsl:656, 732); the lens EMA
    of |v - v_prev| / dt = |g| is |g| (1 - exp(-20 n dt))."""
    sc = _lone()
    v0 = np.array([0.1, 0.05, -0.02], np.float32)
    g = np.array([0.3, -9.81, 0.2], np.float32)
    dt = np.float32(1.0 / 240.0)
    vel = np.zeros((1, 4), np.float32)
    vel[0, :3] = v0
    st = _step(force=tuple(g), dt=float(dt), v_clamp=1e3)
    o = OracleSim(sc.params, precision=precision)
    o.set_state(sc.pos4, vel, vel.copy())  # prev_vel = v0: the first substep's change is dt g too
    n = 12
    for _ in range(n):
        o.substep(st, 5)
    s = o.state_f64()
    x0, v0d, gd, dtd = sc.pos4[0, :3].astype(np.float64), v0.astype(np.float64), g.astype(np.float64), float(dt)
    x = x0 + n * dtd * v0d + dtd * dtd * gd * n * (n + 1) / 2
    v = v0d + n * dtd * gd
    assert np.allclose(s["pos"][0, :3], x, rtol=0, atol=tol * np.abs(x).max()), (s["pos"][0], x)
    assert np.allclose(s["vel"][0, :3], v, rtol=0, atol=tol * np.abs(v).max()), (s["vel"][0], v)
    ema = np.linalg.norm(gd) * (1 - math.exp(-20.0 * n * dtd))
    assert abs(s["vel"][0, 3] - ema) <= max(tol, 1e-6) * ema
    assert s["prev_vel"][0, 3] == 0.0 and s["prev_pressure"][0] == 0.0
    assert abs(s["temperature"][0] - F32(293.15)) <= 1e-5


def test_speed_clamp_lands_on_the_ceiling():
    """Below half rest density (a lone particle) the ceiling is 3 v_clamp; in
    a block at rest density it is v_clamp (sim_solve.wgsl:718-729). The speed
    lands on the ceiling, the direction is kept, and the counter counts the
    particles whose speed exceeded their ceiling."""
    sc = _lone()
    vel = np.zeros((1, 4), np.float32)
    vel[0, :3] = (3.0, -4.0, 0.0)
    vc = 0.5
    o = _run(sc, 64, SUBSTEP, _step(force=(0.0, 0.0, 0.0), dt=1.0 / 240.0, v_clamp=vc), vel=vel)
    v = o.state_f64()["vel"][0, :3]
    assert abs(np.linalg.norm(v) - 3 * vc) < 1e-12
    assert np.allclose(v / np.linalg.norm(v), [0.6, -0.8, 0.0], atol=1e-12)
    assert o.counters()[0] == 1
    blk = fa.scenes.small_block((6, 6, 6), seed=1, spacing=0.01, half=(0.2, 0.2, 0.2))
    rng = np.random.default_rng(3)
    vel = np.zeros((blk.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.6, (blk.n, 3))
    st = _step(force=(0.0, -9.81, 0.0), dt=1.0 / 240.0, v_clamp=vc)
    o = _run(blk, 64, SUBSTEP[:-1], st, vel=vel)
    pre = o.field("work_vel")[:, :3]
    rho = o.field("density")
    ceiling = np.where(rho < 0.5 * _p(blk, 40), 3 * vc, vc)
    over = np.linalg.norm(pre, axis=1) > ceiling
    assert over.sum() > 10 and (~over).sum() > 10
    o.kernel("integrate", st)
    after = np.linalg.norm(o.state_f64()["vel"][:, :3], axis=1)  # resting set, same slots
    assert np.allclose(after[over], ceiling[over], rtol=1e-12)
    assert np.all(after <= ceiling * (1 + 1e-12))
    assert o.counters()[0] == int(over.sum())


def test_wall_clamp_zeroes_only_the_normal_component():
    """A particle that crosses the +x wall within one substep stops on it:
    x = wall_hi = -(box_min + cell) (sim_solve.wgsl:237-239, 735-739), vx = 0
    in the resting velocity and prev_vel, the tangential velocity and motion
    untouched. It starts 2.5 h from the wall, where every wall term is zero
    (a lone particle nearer the wall has its normal velocity removed by the
    divergence solve, the wall acting as fluid at rest)."""
    C0 = _consts(_lone())
    wall_hi = -(C0["bmin"][0] + C0["cell"])
    x0 = wall_hi - 2.5 * C0["h"]
    sc = _lone(pos=(x0, 0.0, 0.0))
    dt = float(np.float32(1.0 / 240.0))
    vel = np.zeros((1, 4), np.float32)
    vel[0, :3] = (10.0, 0.3, -0.2)
    o = _run(sc, 64, SUBSTEP, _step(force=(0.0, 0.0, 0.0), dt=dt, v_clamp=1e3), vel=vel)
    s = o.state_f64()
    assert float(sc.pos4[0, 0]) + dt * 10.0 > wall_hi
    assert s["pos"][0, 0] == wall_hi  # the FP64 oracle's wall: box_min + cell of the f32 parameters, in double
    assert s["vel"][0, 0] == 0.0 and s["prev_vel"][0, 0] == 0.0
    assert s["vel"][0, 1] == np.float64(np.float32(0.3)) and s["vel"][0, 2] == np.float64(np.float32(-0.2))
    assert abs(s["pos"][0, 1] - dt * F32(0.3)) < 1e-15 and abs(s["pos"][0, 2] + dt * F32(0.2)) < 1e-15


def test_temperature_pressure_work_is_isentropic_heating():
    """With a uniform temperature (no conduction) and no viscosity (no viscous
    heating), integrate's temperature is T' = T (1 + beta (p - p_prev) /
    (rho c_p)): the isentropic (dT/dp)_s = T beta / (rho c_p) of water
    (beta = 2.07e-4 / K, c_p = 4184 J / (kg K))."""
    sc = fa.scenes.small_block((6, 6, 6), seed=2, spacing=0.01, half=(0.2, 0.2, 0.2))
    sc.pos4[:, :3] *= np.float32(0.93)  # compressed: the density passes build pressure
    rng = np.random.default_rng(4)
    pp = rng.uniform(0.0, 2000.0, sc.n).astype(np.float32)
    st = _step(force=(0.0, -9.81, 0.0), dt=1.0 / 240.0, v_clamp=1e3)
    o = _run(sc, 64, SUBSTEP[:-1], st, prev_pressure=pp, phys=(0.0, 3000.0, 1.0, 1.0, 6.0))
    p = o.field("pressure")
    rho = o.field("density")
    perm = o.grid()["perm"]
    o.kernel("integrate", st)
    T = o.state_f64()["temperature"]
    T0 = 293.15  # the FP64 oracle's starting temperature (render.rs:1540-1555), in double
    want = T0 * (1 + EXPANSION * (p - pp[perm].astype(np.float64)) / (rho * HEAT_CAPACITY))
    assert np.abs(p - pp[perm]).max() > 100
    assert np.allclose(T, want, rtol=1e-13, atol=0)


# ---- forces_den_apply ---------------------------------------------------------

@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 2e-6)])
def test_rotating_frame_forces_are_the_closed_form(precision, tol):
    """A lone particle (no pressure, no pair terms, far from walls) in a frame
    turning at omega with angular acceleration domega: v' = v + dt (g -
    domega x r - omega x (omega x r) - 2 omega x v), with the centrifugal term
    written as omega^2 r - (omega . r) omega."""
    r = np.array([0.03, 0.01, -0.02])
    sc = _lone(pos=tuple(r))
    v = np.array([0.2, -0.1, 0.05], np.float32)
    om = np.array([0.3, -1.2, 0.9], np.float32)
    dom = np.array([2.0, 0.5, -3.0], np.float32)
    g = np.array([0.5, -9.81, 0.1], np.float32)
    dt = np.float32(1.0 / 240.0)
    vel = np.zeros((1, 4), np.float32)
    vel[0, :3] = v
    st = _step(force=tuple(g), omega=tuple(om), domega=tuple(dom), dt=float(dt), v_clamp=1e3)
    o = _run(sc, precision, SUBSTEP[:5], st, vel=vel)
    got = o.field("work_vel")[0, :3]
    rr = sc.pos4[0, :3].astype(np.float64)
    omd, domd, vd = om.astype(np.float64), dom.astype(np.float64), v.astype(np.float64)
    centrifugal = (omd @ omd) * rr - (omd @ rr) * omd
    want = vd + float(dt) * (g.astype(np.float64) - np.cross(domd, rr) + centrifugal - 2 * np.cross(omd, vd))
    assert np.abs(got - want).max() <= tol * np.abs(want).max(), (got, want)


@pytest.mark.parametrize("faithful", [False, True])
@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 3e-6)])
def test_fingers_blend_toward_the_weighted_finger_velocity(precision, tol, faithful):
    """Fingers within touch_r weigh (1 - q^2)^2 (q = distance / touch_r); the
    particle's screen-plane velocity moves toward their weighted mean velocity
    by 1 - exp(-40 sum(w) dt); z and fingers out of reach do nothing
    (sim_solve.wgsl:103-127)."""
    r = np.array([0.03, 0.01, -0.02], np.float32)
    sc = _lone(pos=tuple(r))
    v = np.array([0.2, -0.1, 0.05], np.float32)
    R = np.float32(0.05)
    fingers = np.array([[r[0], r[1], 0.5, -0.3],                  # q = 0: w = 1
                        [r[0] + 0.5 * R, r[1], -0.2, 0.4],        # q^2 = 0.25
                        [r[0], r[1] + 0.8 * R, 0.1, 0.1],         # q^2 = 0.64
                        [r[0] + 1.2 * R, r[1], 9.0, 9.0]], np.float32)  # out of reach
    dt = np.float32(1.0 / 240.0)
    vel = np.zeros((1, 4), np.float32)
    vel[0, :3] = v
    st = fa.pack_step(force=(0.0, 0.0, 0.0), dt=float(dt), v_clamp=1e3, touches=fingers, touch_n=4, touch_r=float(R))
    o = _run(sc, precision, SUBSTEP[:5], st, vel=vel, faithful=faithful)
    got = o.field("work_vel")[0, :3]
    f = fingers.astype(np.float64)
    q2 = ((f[:, 0] - float(r[0])) ** 2 + (f[:, 1] - float(r[1])) ** 2) / float(R) ** 2
    w = np.where(q2 < 1, (1 - q2) ** 2, 0.0)
    pull = (w[:, None] * f[:, 2:]).sum(0) / w.sum()
    b = 1 - math.exp(-40.0 * w.sum() * float(dt))
    vd = v.astype(np.float64)
    want = np.array([vd[0] * (1 - b) + pull[0] * b, vd[1] * (1 - b) + pull[1] * b, vd[2]])
    assert w[3] == 0 and abs(w[1] - 0.5625) < 1e-6
    assert np.abs(got - want).max() <= tol, (got, want)


# ---- forces_eval ---------------------------------------------------------------

def _cohesion_spline(r, c):
    """Akinci et al. 2013, eq. 4: C(r) = 32/(pi c^9) x {(c - r)^3 r^3 for c/2 <
    r <= c; 2 (c - r)^3 r^3 - c^6/64 for r <= c/2}."""
    k = 32.0 / (PI_REF * c ** 9)
    a = (c - r) ** 3 * r ** 3
    return np.where((r >= c) | (r < 1e-9), 0.0, np.where(2 * r > c, k * a, k * (2 * a - c ** 6 / 64.0)))


def _forces_brute(sc, o, temp_work):
    """float64 O(N^2) forces from the published formulas, given the oracle's
    density and near-density (pinned in test_oracle_pinning.py)."""
    C = _consts(sc)
    pos = o.field("work_pos")[:, :3]
    vel = o.field("work_vel")[:, :3]
    rho = o.field("density")
    rho_n = o.field("rho_near")
    x, r, q, W, gradW, inside = _pairs(C, pos)
    m, h, rho0 = C["m"], C["h"], C["rho0"]
    v_ij = vel[:, None, :] - vel[None, :, :]
    # Morris et al. 1997: a_i = sum_j m (mu_i + mu_j) / (rho_i rho_j) (x_ij . grad W) / (r^2 + 0.01 h^2) v_ij
    F = np.where(inside, (x * gradW).sum(-1) / (r * r + 0.01 * h * h), 0.0)
    pair = m / (rho[:, None] * rho[None, :]) * F
    visc = ((2 * MU * pair)[..., None] * v_ij).sum(1)
    # Monaghan 1989 XSPH: -sum_j m / rho_j W_ij v_ij
    xsph = -((m / rho[None, :] * W)[..., None] * v_ij).sum(1)
    # Clavet et al. 2005 near pressure: sum_j (rho~_i + rho~_j) (1 - q)^2 x_ij / r
    with np.errstate(invalid="ignore", divide="ignore"):
        xhat = np.where(inside[..., None], x / r[..., None], 0.0)
    near = ((rho_n[:, None] + rho_n[None, :]) * np.maximum(1 - q, 0) ** 2)[..., None] * xhat
    near = np.where(inside[..., None], near, 0.0).sum(1)
    # Akinci et al. 2013 cohesion with the symmetrising density correction
    c = 2 * h
    corr = 2 * rho0 / (rho[:, None] + rho[None, :])
    tension = -((corr * m * np.where(inside, _cohesion_spline(r, c), 0.0))[..., None] * xhat).sum(1)
    gamma = SIGMA_WATER * 7040.0 / (21.0 * LATTICE_SIGMA * rho0 * rho0 * c * c)
    # the literals 7040, 21 and 1558.4 (sim_solve.wgsl:163, 170) are exact at FP64
    beta = SIGMA_WATER * (1.0 + COS_CONTACT) * 1558.4 / (LATTICE_SIGMA * rho0 * rho0 * c * c)
    # Akinci et al. 2013 wall adhesion: beta rho0 times the wall kernel's integral, toward each wall in reach
    L = olib()
    lo = C["bmin"] + C["cell"]
    hi = -lo
    adh = np.zeros_like(pos)
    for a in range(3):
        adh[:, a] -= np.array([L.oracle_wall_adhesion(t) for t in (pos[:, a] - lo[a]) / h])
        adh[:, a] += np.array([L.oracle_wall_adhesion(t) for t in (hi[a] - pos[:, a]) / h])
    adh *= beta * rho0
    # Cleary & Monaghan 1999: dT/dt = sum_j 4 kappa^2 / (2 kappa c_p) m / (rho_i rho_j) F (T_i - T_j), and half
    # the pair's viscous dissipation as heat
    dT = temp_work[:, None] - temp_work[None, :]
    heat = (2 * CONDUCTIVITY / HEAT_CAPACITY * pair * dT).sum(1)
    heat -= (0.5 * MU / HEAT_CAPACITY * pair * (v_ij ** 2).sum(-1)).sum(1)
    accel = visc + K_NEAR * near + gamma * tension + adh
    return dict(accel=accel, heat=heat, xsph=xsph, visc=visc, near=near, tension=tension, adh=adh)


def _forces_scene(seed=6):
    sc = fa.scenes.small_block((7, 6, 5), seed=seed, spacing=0.01, half=(0.05, 0.05, 0.05), place="corner")
    rng = np.random.default_rng(seed)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.2, (sc.n, 3))
    temp = rng.uniform(280.0, 320.0, sc.n).astype(np.float32)
    pp = rng.uniform(0.0, 500.0, sc.n).astype(np.float32)
    return sc, vel, temp, pp


@pytest.mark.parametrize("faithful", [False, True])
@pytest.mark.parametrize("precision,tol", [(64, 1e-11), (32, 2e-5)])
def test_pair_forces_match_the_published_formulas(precision, tol, faithful):
    """A block in a corner (walls on three sides: adhesion acts) with random
    velocities and temperatures: every term of forces_eval, and the warm
    start kd = p_prev / 2 / rho^2, p = p_prev / 2 (Bender & Koschier 2015).
    Both the kernels' order and the line-faithful one."""
    sc, vel, temp, pp = _forces_scene()
    st = _step(force=(0.0, -9.81, 0.0), dt=1.0 / 240.0, v_clamp=1e3)
    o = _run(sc, precision, ("grid", "density_div", "forces_eval"), st, vel=vel, temperature=temp,
             prev_pressure=pp, faithful=faithful)
    perm = o.grid()["perm"]
    b = _forces_brute(sc, o, temp[perm].astype(np.float64))
    acc = o.field("accel")
    for name, got, want in (("accel", acc[:, :3], b["accel"]), ("heat", acc[:, 3], b["heat"]),
                            ("xsph", o.field("xsph")[:, :3], b["xsph"])):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < tol, (name, err)
    rho = o.field("density")
    half = 0.5 * pp[perm].astype(np.float64)
    assert np.allclose(o.field("pressure"), half, rtol=1e-7 if precision == 32 else 1e-15)
    assert np.allclose(o.field("kd"), half / rho ** 2, rtol=2e-6 if precision == 32 else 1e-14)


@pytest.mark.parametrize("term,phys", [("visc", (MU, 0.0, 0.0, 0.0, 6.0)), ("near", (0.0, 3000.0, 0.0, 0.0, 6.0)),
                                       ("tension", (0.0, 0.0, 1.0, 0.0, 6.0)), ("adh", (0.0, 0.0, 0.0, 1.0, 6.0))])
def test_each_force_term_alone_matches_its_formula(term, phys):
    """Each term switched on alone (the physics coefficients, sim_solve.wgsl:
    135-137, 161-172), FP64 to 1e-12: a wrong factor in a term that is small
    in the total (viscosity is ~1e-3 of it here) still fails."""
    sc, vel, temp, pp = _forces_scene()
    st = _step(force=(0.0, -9.81, 0.0), dt=1.0 / 240.0, v_clamp=1e3)
    o = _run(sc, 64, ("grid", "density_div", "forces_eval"), st, vel=vel, temperature=temp, prev_pressure=pp,
             phys=phys)
    b = _forces_brute(sc, o, temp[o.grid()["perm"]].astype(np.float64))
    C = _consts(sc)
    c = 2 * C["h"]
    scale = {"visc": 1.0, "near": K_NEAR, "adh": 1.0,
             "tension": SIGMA_WATER * 7040.0 / (21.0 * LATTICE_SIGMA * C["rho0"] ** 2 * c * c)}[term]
    want = scale * b[term]
    got = o.field("accel")[:, :3]
    assert np.linalg.norm(want) > 0
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


def test_conduction_conserves_heat_and_viscosity_dissipates():
    """Away from the walls, with equal masses, the pair conduction terms are
    antisymmetric: sum_i m c_p dT_i/dt = 0, and heat flows down the gradient
    (sum_i T_i dT_i/dt < 0); Morris viscosity removes kinetic energy
    (sum_i m v_i . a_i < 0). Viscous heating switched off (mu = 0) for the
    first, near pressure, cohesion and adhesion off for the second."""
    sc = fa.scenes.small_block((6, 6, 6), seed=8, spacing=0.01, half=(0.5, 0.5, 0.5))
    sc.pos4[:, 1] += np.float32(0.45)
    rng = np.random.default_rng(8)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.2, (sc.n, 3))
    temp = rng.uniform(280.0, 320.0, sc.n).astype(np.float32)
    st = _step(dt=1.0 / 240.0, v_clamp=1e3)
    o = _run(sc, 64, ("grid", "density_div", "forces_eval"), st, vel=vel, temperature=temp,
             phys=(0.0, 3000.0, 1.0, 1.0, 6.0))
    heat = o.field("accel")[:, 3]
    T = temp[o.grid()["perm"]].astype(np.float64)
    assert np.abs(heat).sum() > 0
    assert abs(heat.sum()) < 1e-12 * np.abs(heat).sum()
    assert (T * heat).sum() < 0
    o = _run(sc, 64, ("grid", "density_div", "forces_eval"), st, vel=vel, temperature=temp,
             phys=(MU, 0.0, 0.0, 0.0, 6.0))
    a = o.field("accel")[:, :3]
    v = o.field("work_vel")[:, :3]
    assert (v * a).sum() < 0


# ---- the pressure solves ------------------------------------------------------

def _kappa_scene():
    sc = fa.scenes.small_block((6, 5, 5), seed=9, spacing=0.01, half=(0.05, 0.05, 0.05), place="corner")
    sc.pos4[:, :3] = (sc.pos4[:, :3] - sc.pos4[:, :3].min(0)) * np.float32(0.94) + sc.pos4[:, :3].min(0)
    rng = np.random.default_rng(9)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.1, (sc.n, 3))
    return sc, vel


def _apply_brute(C, pos, vel, kd, wg, dt):
    """Bender & Koschier 2015: v_i -= dt sum_j m (kappa_i/rho_i + kappa_j/rho_j) grad W_ij, plus the static wall's
    kappa_i/rho_i term (the wall joins only the particle's own side)."""
    _, _, _, _, gradW, _ = _pairs(C, pos)
    s = (C["m"] * (kd[:, None] + kd[None, :]))[..., None] * gradW
    return vel - dt * (s.sum(1) + kd[:, None] * wg)


@pytest.mark.parametrize("faithful", [False, True])
@pytest.mark.parametrize("precision,tol", [(64, 1e-11), (32, 1e-5)])
def test_pressure_apply_is_the_dfsph_update(precision, tol, faithful):
    sc, vel = _kappa_scene()
    dt = float(np.float32(1.0 / 240.0))
    st = _step(force=(0.0, -9.81, 0.0), dt=dt, v_clamp=1e3)
    o = _run(sc, precision, ("grid", "density_div"), st, vel=vel, faithful=faithful)
    C = _consts(sc)
    pos = o.field("work_pos")[:, :3]
    before = o.field("work_vel")[:, :3]
    kd = o.field("kd")
    wg = o.field("wall_grad")[:, :3]
    assert np.abs(wg).max() > 0 and (kd > 0).sum() > 10
    o.kernel("div_apply", st)
    want = _apply_brute(C, pos, before, kd, wg, dt)
    got = o.field("work_vel")[:, :3]
    err = np.linalg.norm((got - before) - (want - before)) / np.linalg.norm(want - before)
    assert err < tol, err


@pytest.mark.parametrize("precision,tol", [(64, 1e-11), (32, 1e-5)])
def test_density_pass_is_the_dfsph_prediction_and_its_error_falls(precision, tol):
    """Each den_kappa pass predicts rho* = rho + dt D rho / Dt (D rho / Dt =
    sum_j m (v_i - v_j) . grad W_ij + v_i . wall_grad), sets kappa = max(rho*
    - rho0, 0) alpha / dt^2, kd = kappa / rho, and adds kappa rho to the
    pressure (Bender & Koschier 2015, alg. 3); on a block compressed by 6%
    the mean predicted error falls pass by pass."""
    sc, vel = _kappa_scene()
    dt = float(np.float32(1.0 / 240.0))
    st = _step(force=(0.0, -9.81, 0.0), dt=dt, v_clamp=1e3)
    o = _run(sc, precision, SUBSTEP[:5], st, vel=vel)
    C = _consts(sc)
    pos = o.field("work_pos")[:, :3]
    _, _, _, _, gradW, _ = _pairs(C, pos)
    rho, alpha, wg = o.field("density"), o.field("alpha"), o.field("wall_grad")[:, :3]
    errors = []
    for it in range(8):
        v = o.field("work_vel")[:, :3]
        p_before = o.field("pressure")
        o.kernel("den_kappa", st)
        drho = C["m"] * np.einsum("ijk,ijk->i", v[:, None, :] - v[None, :, :], gradW) + (v * wg).sum(1)
        excess = np.maximum(rho + dt * drho - C["rho0"], 0.0)
        kappa = excess * alpha / (dt * dt)
        # passes past the third leave a few particles a hair above rest density,
        # where rho* - rho0 cancels to the rounding of rho*: values held on the first three
        for name, got, want in (("kd", o.field("kd"), kappa / rho),
                                ("pressure", o.field("pressure") - p_before, kappa * rho)):
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert it >= 3 or err < tol, (it, name, err)
        errors.append(excess.mean() / C["rho0"])
        o.kernel("den_apply", st)
    assert errors[0] > 1e-2
    assert all(b < a for a, b in zip(errors, errors[1:])), errors
    assert errors[-1] < 1e-8, errors


# ---- reduce_stats --------------------------------------------------------------

@pytest.mark.parametrize("faithful", [False, True])
def test_reduce_stats_is_numpy_over_the_fields(faithful):
    sc = fa.scenes.small_block((8, 8, 8), seed=10)
    n, dt, refine, vc = sc.schedule()
    o = OracleSim(sc.params, faithful=faithful)
    o.set_state(sc.pos4)
    for k in range(3):
        o.substep(sc.step_block(k * dt, dt, vc), refine)
    got = o.reduce_stats()
    d = o.field("density").astype(np.float32)
    s = o.state()
    c = d / np.float32(_p(sc, 40)) - np.float32(1.0)
    v = s["vel"]
    sp = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    want = [np.maximum(c.astype(np.float64), 0).mean(), max(float(c.max()), 0.0), d.min(), d.max(),
            s["prev_vel"][:, 3].min(), s["prev_vel"][:, 3].max(), sp.max(), s["temperature"].min(),
            s["temperature"].max(), o.counters()[0], o.counters()[1], sp.min(), v[:, 3].min(), v[:, 3].max()]
    assert abs(got[0] - want[0]) <= 1e-5 * want[0]
    assert np.array_equal(got[1:].astype(np.float64), np.array(want[1:], np.float64)), (got, want)
    if faithful:  # the shader's 256-strided partials, then the halving tree (sim_solve.wgsl:767-778, 807-815)
        c32 = np.maximum(c, np.float32(0))
        part = np.zeros(256, np.float32)
        for l in range(256):
            for v_ in c32[l::256]:
                part[l] = np.float32(part[l] + v_)
        off = 128
        while off:
            part[:off] = part[:off] + part[off:2 * off]
            off //= 2
        assert got[0] == np.float32(part[0] / np.float32(d.size))


# ---- the float-valued dual oracle ---------------------------------------------

def test_fp32_dual_run_takes_the_fp32_path_and_its_tangent_is_the_derivative():
    """Precision 96 (float values, double derivatives) gives every FP32 state
    word; its tangent along initial velocities equals the FP64 dual's over a
    short horizon and the FP32 step's own central difference."""
    sc = fa.scenes.fill_block("dual", 1500, 0.3, seed=3)
    n, dt, refine, vc = sc.schedule()
    rng = np.random.default_rng(3)
    v0 = np.zeros((sc.n, 4), np.float32)
    v0[:, :3] = rng.normal(0, 0.05, (sc.n, 3))
    u = np.zeros((sc.n, 4))
    u[:, :3] = rng.normal(0, 1, (sc.n, 3))
    st = sc.step_block(0.0, dt, vc)
    runs = {}
    for prec in (32, 96, 128):
        o = OracleSim(sc.params, precision=prec)
        o.set_state(sc.pos4, v0)
        if prec != 32:
            o.set_tangent("vel", u)
        for _ in range(3):
            o.substep(st, refine)
        runs[prec] = (o.state(), None if prec == 32 else o.tangent("pos"))
    for k in ("pos", "vel", "prev_vel", "prev_pressure", "temperature"):
        assert np.array_equal(runs[32][0][k].view(np.uint32), runs[96][0][k].view(np.uint32)), k
    a, b = runs[96][1], runs[128][1]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4
