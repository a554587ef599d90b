"""Pins the CPU oracle before any parity claim rests on it.

The reference ships no golden vectors for the step and cannot run here
(Rust + WGSL on Metal), so each oracle piece is checked against what the
mathematics fixes, chosen so a plausible mistake (a dropped term, a wrong
sign, index or axis) fails:
  * wall and wedge polynomials, adhesion: quadrature (sim.rs:419-446,
    594-662, 708-745), and the lattice-gap closures (sim.rs:453-483, 664-696);
  * grid: numpy's float32 division, bincount, cumsum and stable argsort;
  * neighbour list: brute-force pair search, in the stencil order;
  * density, near-density, alpha, divergence kappa: O(N^2) brute force in
    float64 from the closed-form kernel and its derivative;
  * wall gradient: rho0 times the finite-difference gradient of the fill;
  * pressure apply and pair forces: momentum conservation away from walls;
  * tracers: the published PCG hash, IEEE half rounding, a uniform field.
"""
import math

import numpy as np
import pytest

import fluid_amd as fa
from oracle import OracleSim, lib as olib

PI = math.pi


def w_np(q):
    """Cubic spline in h units, sigma = 1/pi (sim.rs:49-60 with h = 1)."""
    q = np.asarray(q, np.float64)
    out = np.where(q < 1, 1 - 1.5 * q * q * (1 - 0.5 * q), np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))
    return out / PI


def dw_np(q):
    """dW/dq of w_np: the derivative the solver's gradient uses (sim_solve.wgsl:184-196)."""
    q = np.asarray(q, np.float64)
    return np.where(q < 1, -3 * q + 2.25 * q * q, np.where(q < 2, -0.75 * (2 - q) ** 2, 0.0)) / PI


def test_kernel_derivative_is_the_derivative():
    q = np.linspace(0.01, 1.99, 397)
    eps = 1e-6
    fd = (w_np(q + eps) - w_np(q - eps)) / (2 * eps)
    assert np.max(np.abs(fd - dw_np(q))) < 1e-7


# ---- wall integrals ------------------------------------------------------

def wall_quadrature(d, gradient):
    n = 20000
    dr = (2.0 - d) / n
    r = d + (np.arange(n) + 0.5) * dr
    cap = 1.0 if gradient else (r - d)
    return float(np.sum(w_np(r) * 2 * PI * r * cap * dr))


def test_wall_integrals_match_quadrature():  # sim.rs:431-446
    L = olib()
    for i in range(40):
        d = i * 0.05
        assert abs(L.oracle_wall_density(d) - wall_quadrature(d, False)) < 1e-4, d
        assert abs(L.oracle_wall_gradient(d) - wall_quadrature(d, True)) < 1e-4, d
        assert abs(L.oracle_wall_density_f32(d) - L.oracle_wall_density(d)) < 1e-6
    assert L.oracle_wall_density(2.0) == 0.0 and L.oracle_wall_gradient(2.0) == 0.0


def test_wall_gradient_is_minus_the_density_slope():
    L = olib()
    for t in np.linspace(0.05, 1.95, 39):
        fd = (L.oracle_wall_density(t + 1e-6) - L.oracle_wall_density(t - 1e-6)) / 2e-6
        assert abs(fd + L.oracle_wall_gradient(t)) < 1e-6, t


def _line_table():
    rho = np.linspace(0.0, 2.0, 4001)
    n = 1000
    dq = 2.0 / n
    z0 = np.arange(n) * dq
    vals = (w_np(np.sqrt(rho[:, None] ** 2 + z0[None, :] ** 2)) +
            w_np(np.sqrt(rho[:, None] ** 2 + (z0[None, :] + dq) ** 2))) / 2 * dq
    return rho, 2.0 * vals.sum(1)


def test_wedge_polynomials_match_quadrature():  # sim.rs:594-662
    L = olib()
    rho_t, line_t = _line_table()
    line = lambda r: np.interp(r, rho_t, line_t, right=0.0)  # noqa: E731
    m = 400
    du = 2.0 / m
    for i in range(7):
        for j in range(7):
            a, b = i * 0.3, j * 0.3
            if a * a + b * b >= 4.0:
                continue
            u1 = a + (np.arange(m) + 0.5) * du
            u2 = b + (np.arange(m) + 0.5) * du
            u1 = u1[u1 < 2.0]
            u2 = u2[u2 < 2.0]
            i2 = float(line(np.sqrt(u1[:, None] ** 2 + u2[None, :] ** 2)).sum() * du * du)
            d2 = -float(line(np.sqrt(a * a + u2 ** 2)).sum() * du)
            assert abs(L.oracle_wedge(a, b) - i2) < 1.5e-3, (a, b)
            assert abs(L.oracle_wedge_d(a, b) - d2) < 8e-3, (a, b)


def test_wall_adhesion_polynomial_matches_the_kernel_quadrature():  # sim.rs:708-745
    L = olib()
    h = 1.2 * 0.0025
    c = 2.0 * h
    n = 4000
    dr = c / n
    r = np.arange(n + 1) * dr
    a = np.where((2 * r < c) | (r > c), 0.0,
                 0.007 / c ** 3.25 * np.maximum(-4 * r * r / c + 6 * r - 2 * c, 0.0) ** 0.25)
    int_a = np.zeros(n + 1)
    for i in range(n - 1, -1, -1):
        int_a[i] = int_a[i + 1] + (a[i] + a[i + 1]) / 2 * dr
    int_j = np.zeros(n + 1)
    for i in range(n - 1, -1, -1):
        int_j[i] = int_j[i + 1] + (i * dr * int_a[i] + (i + 1) * dr * int_a[i + 1]) / 2 * dr
    j0 = 2 * PI * int_j[0]
    for step in range(41):
        d = c * step / 40
        quad = 2 * PI * int_j[int(round(d / dr))]
        assert abs(L.oracle_wall_adhesion(d / h) - quad) < 0.005 * j0, d


def test_the_wall_term_closes_the_half_lattice_gap():  # sim.rs:453-483
    L = olib()
    d = 0.0025
    h = 1.2 * d
    mass = 998.2 * d ** 3
    for layer in range(4):
        z0 = (layer + 0.5) * d
        x, y, z = np.meshgrid(np.arange(-3, 4), np.arange(-3, 4), np.arange(-6, 7), indexing="ij")
        dz = (z + 0.5) * d - z0
        r = np.sqrt((x * d) ** 2 + (y * d) ** 2 + dz ** 2)
        w = mass * w_np(r / h) / h ** 3
        bulk, half = w.sum(), w[z >= 0].sum()
        err = (half + 998.2 * L.oracle_wall_density(z0 / h) - bulk) / bulk
        assert abs(err) < 0.03, layer


def test_the_wedge_term_closes_the_edge_lattice_gap():  # sim.rs:664-696
    L = olib()
    d = 0.0025
    h = 1.2 * d
    mass = 998.2 * d ** 3
    for lx, ly in [(0, 0), (0, 1), (1, 1), (0, 2), (2, 2)]:
        x0, y0 = (lx + 0.5) * d, (ly + 0.5) * d
        x, y, z = np.meshgrid(np.arange(-6, 7), np.arange(-6, 7), np.arange(-6, 7), indexing="ij")
        r = np.sqrt(((x + 0.5) * d - x0) ** 2 + ((y + 0.5) * d - y0) ** 2 + (z * d) ** 2)
        w = mass * w_np(r / h) / h ** 3
        quarter = w[(x >= 0) & (y >= 0)].sum()
        fill = L.oracle_wall_density(x0 / h) + L.oracle_wall_density(y0 / h) - L.oracle_wedge(x0 / h, y0 / h)
        err = (quarter + 998.2 * fill - w.sum()) / w.sum()
        assert abs(err) < 0.06, (lx, ly)


# ---- scenes through the oracle --------------------------------------------

def _tiny(lattice=(6, 5, 4), seed=0, half=(0.05, 0.05, 0.05), place="floor", h_factor=1.2):
    return fa.scenes.small_block(lattice, seed=seed, spacing=0.01, half=half, place=place, h_factor=h_factor)


def _oracle(scene, precision, vel=None, dt=1.0 / 240.0, cap=96, faithful=False):
    o = OracleSim(scene.params, nbr_cap=cap, precision=precision, faithful=faithful)
    o.set_state(scene.pos4, vel)
    step = fa.pack_step(force=(0.0, -9.81, 0.0), dt=dt, v_clamp=1e30)
    return o, step


def test_grid_is_numpy_counting_sort():
    sc = fa.scenes.dam_break(seed=3)
    o, step = _oracle(sc, 32)
    o.kernel("grid", step)
    g = o.grid()
    p = sc.pos4[:, :3].astype(np.float32)
    mn = np.array(sc.grid.min, np.float32)
    cell = np.float32(sc.grid.cell)
    dims = np.array(sc.grid.dims, np.uint32)
    q = (p - mn) / cell  # float32 IEEE division, as the shader
    coord = np.minimum(np.where(q > 0, np.floor(q), 0).astype(np.uint32), dims - 1)
    cid = (coord[:, 2] * dims[1] + coord[:, 1]) * dims[0] + coord[:, 0]
    assert np.array_equal(g["cellid"], cid)
    counts = np.bincount(cid, minlength=sc.grid.cell_count())
    assert np.array_equal(g["counts"], counts)
    assert np.array_equal(g["starts"], np.concatenate([[0], np.cumsum(counts)]))
    assert np.array_equal(g["perm"], np.argsort(cid, kind="stable"))


def _brute(scene, o, vel_slot, dt):
    """float64 O(N^2) density, near-density, alpha and divergence kappa."""
    pos = o.field("work_pos")[:, :3]
    # the method's constants as the SimParams block carries them (f32)
    f = lambda off: float(np.frombuffer(scene.params, np.float32, 1, off)[0])  # noqa: E731
    hh, m, rho0, inv_h, sig, sig_h, sup2 = f(32), f(36), f(40), f(44), f(48), f(52), f(56)
    x = pos[:, None, :] - pos[None, :, :]
    r2 = (x ** 2).sum(-1)
    r = np.sqrt(r2)
    q = r * inv_h
    inside = (r2 < sup2) & ~np.eye(len(pos), dtype=bool)
    W = np.where(inside, w_np(q) * PI * sig, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        g = np.where(inside[..., None], (dw_np(q) * PI * sig_h / r)[..., None] * x, 0.0)
    L = olib()
    bmin = np.frombuffer(scene.params, np.float32, 3, 0).astype(np.float64)
    cell = float(np.frombuffer(scene.params, np.float32, 1, 12)[0])
    lo = bmin + cell
    hi = -lo
    tl = (pos - lo) * inv_h
    th = (hi - pos) * inv_h
    fill = np.zeros(len(pos))
    for a in range(3):
        fill += np.array([L.oracle_wall_density(t) for t in tl[:, a]])
        fill += np.array([L.oracle_wall_density(t) for t in th[:, a]])
    for (a, b) in [(0, 1), (0, 2), (1, 2)]:
        for A in (tl, th):
            for B in (tl, th):
                fill -= np.array([L.oracle_wedge(u, v) for u, v in zip(A[:, a], B[:, b])])
    rho = m * (sig + W.sum(1)) + rho0 * fill
    rn = np.where(inside, np.maximum(1 - q, 0) ** 3, 0.0).sum(1)
    wg = o.field("wall_grad")[:, :3]
    gs = m * g.sum(1) + wg
    gsq = m * m * (g ** 2).sum(-1).sum(1)
    alpha = rho / np.maximum((gs ** 2).sum(1) + gsq, 1e-4)
    v = vel_slot[:, :3]
    drho = m * np.einsum("ijk,ijk->i", v[:, None, :] - v[None, :, :], g) + (v * wg).sum(1)
    kd = np.maximum(drho, 0) * alpha / max(dt, 1e-9) / rho
    return dict(density=rho, rho_near=rn, alpha=alpha, kd=kd, nbr_count=inside.sum(1))


@pytest.mark.parametrize("faithful", [False, True])
@pytest.mark.parametrize("precision,tol", [(64, 1e-11), (32, 1e-5)])
def test_density_sweep_matches_brute_force(precision, tol, faithful):
    sc = _tiny(seed=1)
    rng = np.random.default_rng(5)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.1, (sc.n, 3))
    dt = 1.0 / 240.0
    o, step = _oracle(sc, precision, vel, dt, faithful=faithful)
    o.kernel("grid", step)
    o.kernel("density_div", step)
    perm = o.grid()["perm"]
    b = _brute(sc, o, vel[perm].astype(np.float64), float(np.float32(dt)))
    assert np.array_equal(o.grid()["nbr_n"], b["nbr_count"])
    for k in ("density", "rho_near", "alpha", "kd"):
        got = o.field(k)
        err = np.linalg.norm(got - b[k]) / np.linalg.norm(b[k])
        assert err < tol, (k, err)


def test_a_lone_particle_reads_its_own_mass_plus_the_wall_fill():  # render.rs:4521-4528
    sc = _tiny(lattice=(1, 1, 1), half=(0.2, 0.2, 0.2))
    o, step = _oracle(sc, 64)
    o.density_only(step)
    m = float(np.frombuffer(sc.params, np.float32, 1, 36)[0])
    sig = float(np.frombuffer(sc.params, np.float32, 1, 48)[0])
    inv_h = float(np.frombuffer(sc.params, np.float32, 1, 44)[0])
    rho0 = float(np.frombuffer(sc.params, np.float32, 1, 40)[0])
    y = float(sc.pos4[0, 1])
    bmin = np.frombuffer(sc.params, np.float32, 3, 0).astype(np.float64)
    lo = bmin[1] + float(np.frombuffer(sc.params, np.float32, 1, 12)[0])
    fill = olib().oracle_wall_density((y - lo) * inv_h)
    assert fill > 0.1  # the particle rests on the floor
    assert abs(o.field("density")[0] - (m * sig + rho0 * fill)) < 1e-12 * rho0


def test_wall_gradient_is_rho0_times_the_fill_gradient():
    sc = _tiny(lattice=(5, 5, 5), half=(0.035, 0.035, 0.035), place="corner", h_factor=1.2)
    o, step = _oracle(sc, 64)
    o.kernel("grid", step)
    o.kernel("density_div", step)
    pos = o.field("work_pos")[:, :3]
    wg = o.field("wall_grad")[:, :3]
    L = olib()
    hh = float(np.frombuffer(sc.params, np.float32, 1, 32)[0])
    bmin = np.frombuffer(sc.params, np.float32, 3, 0).astype(np.float64)
    lo = bmin + float(np.frombuffer(sc.params, np.float32, 1, 12)[0])

    def fill(p):
        tl, th = (p - lo) / hh, (-lo - p) / hh
        f = sum(L.oracle_wall_density(tl[a]) + L.oracle_wall_density(th[a]) for a in range(3))
        for (a, b) in [(0, 1), (0, 2), (1, 2)]:
            for A in (tl, th):
                for B in (tl, th):
                    f -= L.oracle_wedge(A[a], B[b])
        return f

    eps = 1e-7
    fd = np.array([[(fill(p + eps * e) - fill(p - eps * e)) / (2 * eps) for e in np.eye(3)] for p in pos])
    ref = 998.2 * fd
    # the wedge_d fit is a separate fit of the wedge slope (1.1% of scale)
    err = np.linalg.norm(wg - ref) / np.linalg.norm(ref)
    assert err < 0.02, err
    assert np.all(np.sign(wg[np.abs(ref) > 0.1 * np.abs(ref).max()]) ==
                  np.sign(ref[np.abs(ref) > 0.1 * np.abs(ref).max()]))


def test_neighbour_list_is_the_brute_force_set_in_stencil_order():
    sc = _tiny(lattice=(7, 6, 5), seed=2)
    o, step = _oracle(sc, 32)
    o.kernel("grid", step)
    o.kernel("density_div", step)
    pos = o.field("work_pos")[:, :3].astype(np.float32)
    nbr, nn = o.nbr(), o.grid()["nbr_n"]
    g = sc.grid
    mn, cell, dims = np.array(g.min, np.float32), np.float32(g.cell), np.array(g.dims)
    coord = np.minimum(np.floor(np.maximum((pos - mn) / cell, 0)).astype(np.int64), dims - 1)
    h2 = np.float32(4.0) * np.float32(float(np.frombuffer(sc.params, np.float32, 1, 32)[0])) ** 2
    h2 = np.frombuffer(sc.params, np.float32, 1, 56)[0]
    for i in range(sc.n):
        x = pos[i] - pos
        r2 = (x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]) + x[:, 2] * x[:, 2]
        js = [j for j in np.nonzero(r2 < h2)[0] if j != i]
        off = coord[js] - coord[i]
        key = (off[:, 2] + 1) * 9 + (off[:, 1] + 1) * 3 + (off[:, 0] + 1)
        expect = [j for _, j in sorted(zip(key, js))]
        assert list(nbr[i, : nn[i]]) == expect, i


def test_faithful_list_is_the_brute_force_set_in_lane_lockstep_order():
    """The line-faithful walk (sim_solve.wgsl:436-466): lane l of 8 takes the
    stencil cells l, l + 8, l + 16, l + 24 (x fastest); the appends of one
    loop iteration land in lane order. So neighbour j of i sorts by (round
    of its cell = cell // 8, its place in its cell, lane = cell % 8)."""
    sc = _tiny(lattice=(7, 6, 5), seed=2)
    o, step = _oracle(sc, 32, faithful=True)
    o.kernel("grid", step)
    o.kernel("density_div", step)
    pos = o.field("work_pos")[:, :3].astype(np.float32)
    nbr, nn, starts = o.nbr(), o.grid()["nbr_n"], o.grid()["starts"]
    g = sc.grid
    mn, cell, dims = np.array(g.min, np.float32), np.float32(g.cell), np.array(g.dims)
    coord = np.minimum(np.floor(np.maximum((pos - mn) / cell, 0)).astype(np.int64), dims - 1)
    cid = (coord[:, 2] * dims[1] + coord[:, 1]) * dims[0] + coord[:, 0]
    h2 = np.frombuffer(sc.params, np.float32, 1, 56)[0]
    for i in range(sc.n):
        x = pos[i] - pos
        r2 = (x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]) + x[:, 2] * x[:, 2]
        js = [j for j in np.nonzero(r2 < h2)[0] if j != i]
        off = coord[js] - coord[i]
        stencil = (off[:, 2] + 1) * 9 + (off[:, 1] + 1) * 3 + (off[:, 0] + 1)
        place = np.array(js) - starts[cid[js]]
        expect = [j for _, _, _, j in sorted(zip(stencil // 8, place, stencil % 8, js))]
        assert list(nbr[i, : nn[i]]) == expect, i


def _interior_scene():
    # a block far from every wall: no wall fill, no wall gradient, no adhesion
    return _tiny(lattice=(6, 6, 6), half=(0.5, 0.5, 0.5), seed=4)


def _shift_to_centre(sc):
    sc.pos4[:, 1] += 0.45
    return sc


def test_pressure_apply_conserves_momentum_away_from_walls():
    sc = _shift_to_centre(_interior_scene())
    rng = np.random.default_rng(1)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.2, (sc.n, 3))
    o, step = _oracle(sc, 64, vel)
    o.kernel("grid", step)
    o.kernel("density_div", step)
    assert np.all(o.field("wall_grad") == 0.0)
    before = o.field("work_vel")[:, :3].copy()
    o.kernel("div_apply", step)
    dv = o.field("work_vel")[:, :3] - before
    assert np.abs(dv).sum() > 0
    assert np.linalg.norm(dv.sum(0)) < 1e-12 * np.abs(dv).sum()


def test_pair_forces_conserve_momentum_away_from_walls():
    sc = _shift_to_centre(_interior_scene())
    rng = np.random.default_rng(2)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.2, (sc.n, 3))
    o, step = _oracle(sc, 64, vel)
    for k in ("grid", "density_div", "forces_eval"):
        o.kernel(k, step)
    acc = o.field("accel")[:, :3]
    assert np.abs(acc).sum() > 0
    assert np.linalg.norm(acc.sum(0)) < 1e-11 * np.abs(acc).sum()


def test_a_substep_with_zero_dt_guard_is_finite():
    sc = _tiny(seed=3)
    o, _ = _oracle(sc, 32)
    o.density_only(fa.pack_step(dt=0.0))
    assert np.all(np.isfinite(o.field("kd")))


# ---- tracers ----------------------------------------------------------------

def pcg_np(x):
    """PCG hash (Jarzynski & Olano 2020), the one sim_tracers.wgsl:54-58 uses."""
    x = np.asarray(x, np.uint64)
    st = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((st >> ((st >> 28) + 4)) ^ st) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word).astype(np.uint32)


def test_pcg_is_the_published_hash():
    xs = np.random.default_rng(0).integers(0, 2 ** 32, 1000, dtype=np.uint64)
    L = olib()
    assert all(L.oracle_pcg(int(x)) == int(pcg_np(x)) for x in xs)


def test_half_conversion_is_ieee_round_to_nearest_even():
    L = olib()
    v = np.concatenate([np.random.default_rng(1).normal(0, 30, 3000), [65504.0, 65520.0, 1e-8, -3e-5]])
    for x in v.astype(np.float32):
        hb = L.oracle_f16(float(x))
        assert hb == int(np.float16(x).view(np.uint16)), x
        assert L.oracle_f16_to_f32(hb) == float(np.float16(x))


def test_tracers_ride_a_uniform_flow():
    sc = _tiny(lattice=(8, 8, 8), half=(0.06, 0.06, 0.06), seed=6)
    o, step = _oracle(sc, 32)
    v = np.array([0.05, -0.02, 0.01], np.float32)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[:, :3] = v
    o.set_state(sc.pos4, vel)
    rng = np.random.default_rng(3)
    ext = (0.06, 0.06)
    pts = rng.uniform([-0.035, -0.055, -0.035], [0.035, -0.03, 0.035], (500, 3))
    tr = np.stack([fa.pack_tracer(p, 0.0, ext) for p in pts])
    dt = 1.0 / 120.0
    out = o.tracers(fa.pack_step(dt=dt, seed=1), tr)
    x0 = ((tr[:, 0] & 0xFFFF) / 65535.0 * 2 - 1) * ext[0]
    x1 = ((out[:, 0] & 0xFFFF) / 65535.0 * 2 - 1) * ext[0]
    moved = np.abs(x1 - x0 - v[0] * dt) < 2 * ext[0] / 65535
    respawned = ~moved
    assert respawned.mean() < 0.02   # recycling rate dt / 3 s plus the cells the block does not cover
    assert moved.sum() > 450


def test_fixed_sequence_exp_is_within_two_ulp():
    L = olib()
    xs = np.concatenate([np.random.default_rng(4).uniform(-80, 80, 20000), [0.0, -1e-8, 1e-8, -87.0, 88.0]])
    xs = xs.astype(np.float32)
    got = np.array([L.oracle_exp_fixed(float(x)) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    ulp = np.abs(got - ref) / np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert ulp.max() <= 2.0, ulp.max()
