"""The benchmark configurations at their own sizes, through the C ABI.

- block_1m (1,048,576 particles, sinusoidal tilt, 4,096 tracers): one whole
  frame of the reference schedule (3 substeps at R = 5, stats, tracers) from
  the seeded block, every state word, grid word, tracer and counter equal to
  the FP32 oracle's; then the bench's developed window (330 substeps), the
  state finite and inside the box with no list overflow, and one more
  substep from that state equal to the oracle's, staged and unstaged.
- block_1m at the configs' tolerances: two substeps with both solves
  iterated, the same passes and words as the oracle's tolerance mode.
- block_8m (8,388,608 particles, random-walk tilt): one substep from the
  seeded block and one from 40 substeps on, each equal to the oracle's.
- ensemble_256 (256 scenes, 2,048 to 65,536 particles each) under hipGraph
  replay: every scene equal to a lone sim, and the smallest, the largest and
  six seeded others equal to the oracle.
- adjoint_64k (65,536 particles, 64 substeps, a checkpoint every 8): the
  gradient's directional derivatives along initial velocities, viscosity and
  the 64 gravity vectors against the forward-mode derivative of the same
  FP32 trajectory (at the bench's 4.17 ms substep, where the step's
  linearisation is unstable), and against the FP64 dual-number oracle and
  central differences with eps = 1e-3 (at a 1.04 ms substep, where it is
  stable).

Integer and float words are compared bit for bit (the HIP step evaluates
the oracle's operations in the oracle's order); the stats block is held to
numpy over the downloaded fields (min and max exactly, the mean compression
to 1e-5). Each test prints what it ran before it runs it.
"""
import math

import numpy as np
import pytest

import fluid_amd as fa
from fluid_amd import adjoint as fad
from oracle import OracleSim

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _say(msg):
    print(msg, flush=True)


def _assert_state_bits(got, ref, what):
    for k in STATE:
        same = _bits(got[k]) == _bits(ref[k])
        assert same.all(), f"{what}: {k} differs in {int((~same).sum())} words"


def _expected_stats(sim, rho0, clamps, overflow):
    """The 14-float stats block (sim_solve.wgsl:819-873) recomputed with numpy
    from the sim's downloaded fields, in float32 operations (min and max are
    exact in any order) and the compression mean in float64."""
    d = sim.field("density")
    st = sim.state()
    c = d / np.float32(rho0) - np.float32(1.0)
    v = st["vel"]
    sp = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    out = np.zeros(14, np.float64)
    out[0] = np.maximum(c.astype(np.float64), 0.0).sum() / d.size
    out[1] = max(float(c.max()), 0.0)
    out[2], out[3] = d.min(), d.max()
    out[4], out[5] = st["prev_vel"][:, 3].min(), st["prev_vel"][:, 3].max()
    out[6], out[11] = sp.max(), sp.min()
    out[7], out[8] = st["temperature"].min(), st["temperature"].max()
    out[9], out[10] = float(np.float32(clamps)), float(np.float32(overflow))
    out[12], out[13] = v[:, 3].min(), v[:, 3].max()
    return out


def _check_stats(got, want):
    assert abs(got[0] - want[0]) <= 1e-5 * max(abs(want[0]), 1e-12), (got[0], want[0])
    assert np.array_equal(got[1:].astype(np.float64), want[1:]), (got, want)


def _check_developed(sc, st, counters, substeps):
    """Physical bounds of a developed state: every word finite, every particle
    inside the walls (sim_solve.wgsl:737-739), no neighbour-list overflow."""
    for k in STATE:
        assert np.isfinite(st[k]).all(), k
    for a in range(3):
        wl = np.float32(np.float32(sc.grid.min[a]) + np.float32(sc.grid.cell))  # box_min + cell
        assert st["pos"][:, a].min() >= wl and st["pos"][:, a].max() <= -wl, a
    clamps, overflow = counters
    assert overflow == 0, f"{overflow} neighbour pairs past the cap"
    _say(f"{sc.name}: {substeps} substeps, speed clamps {clamps} "
         f"({clamps / (sc.n * substeps):.2%} of particle-substeps), no list overflow, "
         f"speed max {np.sqrt((st['vel'][:, :3].astype(np.float64) ** 2).sum(1)).max():.3f} m/s")


def _oracle_from(sc, state, step, refine):
    o = OracleSim(sc.params)
    o.set_state(state["pos"], state["vel"], state["prev_vel"], state["prev_pressure"], state["temperature"])
    o.substep(step, refine)
    out = o.state()
    out["grid"] = o.grid()
    out["counters"] = o.counters()
    o.close()
    return out


@pytest.mark.timeout(600)
def test_block_1m_frame_is_bit_exact_with_the_oracle():
    sc = fa.scenes.block_1m(seed=0)
    nsub, dt, refine, vc = sc.schedule()
    _say(f"block_1m: {sc.n} particles, {len(sc.tracers)} tracers, one frame of {nsub} substeps at R = {refine}")
    steps = [sc.step_block(k * dt, dt, vc) for k in range(nsub)]
    frame = fa.pack_step(force=sc.gravity(0.0), dt=nsub * dt, seed=1)
    with fa.Sim(sc.params, tracers=len(sc.tracers)) as s:
        s.upload_state(sc.pos4)
        s.upload_tracers(sc.tracers)
        for st in steps:
            s.substeps(st, 1, refine)
        s.reduce_stats()
        stats = s.read_stats()
        s.tracers(frame)
        g = s.state()
        grid = s.grid()
        tr = s.download_tracers()
        counters = s.counters()
        want_stats = _expected_stats(s, sc.rho0, *counters)
    _check_stats(stats, want_stats)
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    for st in steps:
        o.substep(st, refine)
    ot = o.tracers(frame, sc.tracers)
    _assert_state_bits(g, o.state(), "block_1m frame")
    og = o.grid()
    for k in ("cellid", "starts", "perm", "nbr_n"):
        assert np.array_equal(grid[k], og[k]), k
    assert counters == o.counters()
    assert np.array_equal(tr.reshape(-1), ot.reshape(-1))
    o.close()


@pytest.mark.timeout(600)
def test_block_1m_developed_flow_stays_bounded_and_bit_exact(monkeypatch):
    """The bench's developed window (30 + 300 substeps of the tilting block,
    by then fallen and compressed, about a tenth of the particles speed
    clamped each substep), then one more substep from that state on the
    oracle, on a sim that stages its tiles and on one that stages none
    (FLUID_LDS_BYTES = 8 KB): the same words everywhere."""
    sc = fa.scenes.block_1m(seed=0)
    _, dt, refine, vc = sc.schedule()
    window = 330
    _say(f"block_1m: {window} substeps at R = {refine}, then one substep three ways")
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        for k in range(window):
            s.substeps(sc.step_block(k * dt, dt, vc), 1, refine)
            if (k + 1) % 110 == 0:
                _check_developed(sc, s.state(), s.counters(), k + 1)
        dev = s.state()
        step = sc.step_block(window * dt, dt, vc)
        c0 = s.counters()
        s.substeps(step, 1, refine)
        staged = s.state()
        staged_nn = s.grid()["nbr_n"]
        c1 = s.counters()
        assert s.describe()["stage_cap"] > 1000
    monkeypatch.setenv("FLUID_LDS_BYTES", "8192")
    with fa.Sim(sc.params) as u:
        assert u.describe()["stage_cap"] < 200
        u.upload_state(dev["pos"], dev["vel"], dev["prev_vel"], dev["prev_pressure"], dev["temperature"])
        u.substeps(step, 1, refine)
        _assert_state_bits(u.state(), staged, "unstaged against staged")
        assert np.array_equal(u.grid()["nbr_n"], staged_nn)
    ref = _oracle_from(sc, dev, step, refine)
    _assert_state_bits(staged, ref, "developed block_1m against the oracle")
    assert np.array_equal(staged_nn, ref["grid"]["nbr_n"])
    assert (c1[0] - c0[0], c1[1] - c0[1]) == ref["counters"]
    _say(f"developed state: neighbours per particle mean {staged_nn.mean():.1f}, max {staged_nn.max()}")


@pytest.mark.timeout(600)
def test_block_1m_at_the_configs_tolerances_is_bit_exact_with_the_oracle():
    """Config 2's tolerances (divergence 0.1%, density 0.01%; caps of 64
    passes each) on block_1m after 30 substeps of the reference schedule:
    two substeps on the GPU and on the oracle's tolerance mode, the same
    pass counts and every state word equal."""
    sc = fa.scenes.block_1m(seed=0)
    _, dt, refine, vc = sc.schedule()
    _say(f"block_1m: 30 substeps at R = {refine}, then 2 in tolerance mode")
    setting = (1e-4, 1, 64, 1e-3, 64)
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        for k in range(30):
            s.substeps(sc.step_block(k * dt, dt, vc), 1, refine)
        dev = s.state()
        s.set_tolerance(*setting)
        got = []
        for k in range(30, 32):
            s.substeps(sc.step_block(k * dt, dt, vc), 1, refine)
            got.append(s.iterations())
        g = s.state()
    o = OracleSim(sc.params)
    o.set_state(dev["pos"], dev["vel"], dev["prev_vel"], dev["prev_pressure"], dev["temperature"])
    o.set_tolerance(*setting)
    want = []
    for k in range(30, 32):
        o.substep(sc.step_block(k * dt, dt, vc), refine)
        want.append(o.passes()[0])
    _say(f"(density, divergence) passes: GPU {got}, oracle {want}")
    assert got == want
    _assert_state_bits(g, o.state(), "block_1m in tolerance mode")
    o.close()


@pytest.mark.timeout(900)
def test_block_8m_substeps_are_bit_exact_with_the_oracle():
    sc = fa.scenes.block_8m(seed=0)
    _, dt, refine, vc = sc.schedule()
    later = 40
    _say(f"block_8m: {sc.n} particles, {sc.grid.cell_count()} cells, R = {refine}; substep 1 and substep {later + 2}")
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        first = sc.step_block(0.0, dt, vc)
        s.substeps(first, 1, refine)
        a = s.state()
        ca = s.counters()
        for k in range(1, later + 1):
            s.substeps(sc.step_block(k * dt, dt, vc), 1, refine)
        b = s.state()
        _check_developed(sc, b, s.counters(), later + 1)
        cb = s.counters()
        step = sc.step_block((later + 1) * dt, dt, vc)
        s.substeps(step, 1, refine)
        c = s.state()
        cc = s.counters()
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    o.substep(first, refine)
    _assert_state_bits(a, o.state(), "block_8m substep 1")
    assert ca == o.counters()
    o.close()
    ref = _oracle_from(sc, b, step, refine)
    _assert_state_bits(c, ref, f"block_8m substep {later + 2}")
    assert (cc[0] - cb[0], cc[1] - cb[1]) == ref["counters"]


def _ensemble_plan(ens, substeps):
    sched = [sc.schedule() for sc in ens]
    plan = []
    for k in range(substeps):
        plan.append(([sc.step_block(k * sd[1], sd[1], sd[3], seed=k) for sc, sd in zip(ens, sched)],
                     [sd[2] for sd in sched]))
    return plan


@pytest.mark.timeout(900)
def test_ensemble_256_under_graph_replay_matches_lone_sims_and_the_oracle():
    ens = fa.scenes.ensemble(256, seed=0)
    substeps = 4
    plan = _ensemble_plan(ens, substeps)
    counts = np.array([sc.n for sc in ens])
    _say(f"ensemble_256: {counts.sum()} particles in 256 scenes ({counts.min()} to {counts.max()}), "
         f"{substeps} graph-replayed substeps, refine counts {sorted(set(plan[0][1]))}")
    with fa.Batch([sc.params for sc in ens]) as b:
        b.use_graph(True)
        for i, sc in enumerate(ens):
            b.upload_state(i, sc.pos4)
        for steps, refine in plan:
            b.substep(steps, refine)
        b.reduce_stats()
        stats = b.read_stats()
        counters = b.counters()
        got = [b.state(i) for i in range(len(ens))]
        perms = [b.grid(i)["perm"] for i in range(len(ens))]
    for i, sc in enumerate(ens):
        with fa.Sim(sc.params) as s:
            s.upload_state(sc.pos4)
            for steps, refine in plan:
                s.substeps(steps[i], 1, refine[i])
            s.reduce_stats()
            _assert_state_bits(got[i], s.state(), f"scene {i} ({sc.n} particles) against a lone sim")
            assert np.array_equal(perms[i], s.grid()["perm"]), i
            assert tuple(int(x) for x in counters[i]) == s.counters(), i
            assert np.array_equal(_bits(stats[i]), _bits(s.read_stats())), i
    rng = np.random.default_rng(1)
    picks = [int(counts.argmin()), int(counts.argmax())]
    picks += [int(x) for x in rng.choice([i for i in range(len(ens)) if i not in picks], 6, replace=False)]
    for i in picks:
        sc = ens[i]
        o = OracleSim(sc.params)
        o.set_state(sc.pos4)
        for steps, refine in plan:
            o.substep(steps[i], refine[i])
        _assert_state_bits(got[i], o.state(), f"scene {i} ({sc.n} particles) against the oracle")
        assert tuple(int(x) for x in counters[i]) == o.counters(), i
        o.close()


# ---- config 5: the adjoint at its own size --------------------------------

MU = 1.002e-3


def _adjoint_config(frame_substeps=None, seed=0, n=65536, steps=64, physics="full"):
    """65,536 particles filling 30% of the unit box on the floor, seeded
    initial velocities (N(0, 0.05) m/s a component), a seeded random-walk
    tilt of gravity (sigma 0.05 rad a substep), and seeded targets: the
    centroid moved by (0.02, 0.01, -0.01) m and the 32-bin density histogram
    of the initial block moved 0.03 m along x. The substep is the reference
    schedule's at this spacing (2 a 1/120 s frame, R = 5), or 1/120 s split
    into `frame_substeps` with R = refine_passes(dt_sub) (render.rs:77-79).
    `physics`: "full", the reference's step, or "viscous" (near pressure,
    cohesion, adhesion and the XSPH blend off)."""
    sc = fa.scenes.fill_block("adjoint_64k", n, 0.3, seed=seed)
    _, dt, refine, vc = sc.schedule()
    if frame_substeps:
        dt = (1.0 / 120.0) / frame_substeps
        refine = fa.refine_passes(dt)
        vc = 0.4 * sc.spacing / dt
    rng = np.random.default_rng(seed + 17)
    v0 = np.zeros((sc.n, 4), np.float32)
    v0[:, :3] = rng.normal(0.0, 0.05, (sc.n, 3))
    th = np.cumsum(rng.normal(0.0, 0.05, steps))
    grav = np.stack([9.81 * np.sin(th), -9.81 * np.cos(th), np.zeros(steps)], 1)
    c_t = sc.pos4[:, :3].astype(np.float64).mean(0) + np.array([0.02, 0.01, -0.01])
    h_t = fad.soft_histogram(sc.pos4[:, :3].astype(np.float64) + np.array([0.03, 0.0, 0.0]), sc.half)
    return dict(sc=sc, dt=dt, refine=refine, vc=vc, v0=v0, grav=grav, c_t=c_t, h_t=h_t,
                physics=fad.FULL if physics == "full" else fad.VISCOUS)


def _loss_of(pos, C):
    return float(fad.loss(np.asarray(pos, np.float64)[:, :3], C["sc"].half, C["c_t"], C["h_t"]))


def _gpu_loss(C, v0=None, grav=None, mu=MU):
    grav = C["grav"] if grav is None else grav
    adj = fad.Adjoint(C["sc"].params, len(grav), len(grav), physics=C["physics"])
    adj.set_mu(mu)
    pos, _ = adj.forward(C["sc"].pos4, C["v0"] if v0 is None else v0, fad.steps_for(grav, C["dt"], C["vc"]),
                         [C["refine"]] * len(grav))
    adj.close()
    return _loss_of(pos, C)


def _gpu_gradient(C, every):
    """The GPU adjoint's gradient of the loss (through the torch.autograd
    wrapper) with respect to v0, mu and the gravity vectors."""
    import torch
    sc, steps = C["sc"], len(C["grav"])
    adj = fad.Adjoint(sc.params, steps, every, physics=C["physics"])
    Sim = fad.simulate_function()
    v = torch.tensor(C["v0"][:, :3], dtype=torch.float32, requires_grad=True)
    m = torch.tensor(MU, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(C["grav"], dtype=torch.float64, requires_grad=True)
    x = Sim.apply(v, m, g, adj, sc.pos4, C["dt"], C["vc"], C["refine"]).double()
    L = fad.loss(x, sc.half, C["c_t"], torch.tensor(C["h_t"]), xp=torch)
    L.backward()
    fwd_ms, bwd_ms = adj.timing()
    adj.close()
    out = dict(v0=v.grad.numpy().astype(np.float64), mu=float(m.grad), g=g.grad.numpy())
    assert np.isfinite(out["v0"]).all() and np.isfinite(out["g"]).all() and math.isfinite(out["mu"])
    _say(f"loss {float(L.detach()):.6e}; forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms "
         f"({bwd_ms / fwd_ms:.2f}x)")
    return out


def _oracle_run(C, precision, v0=None, grav=None, mu=MU, dv=None, dmu=0.0, dg=None):
    """The oracle's loss at `precision` (32, 64; 96 or 128 with a seed) and,
    with a seed (dv, dmu, dg), the loss's directional derivative along it:
    precision 128 is the FP64 step's derivative (dual numbers over double),
    96 the FP32 step's own (float values, double derivatives: the trajectory
    the GPU runs, word for word)."""
    import torch
    sc = C["sc"]
    grav = C["grav"] if grav is None else grav
    o = OracleSim(sc.params, precision=precision)
    ph = C["physics"]
    o.set_physics(mu, ph["k_near"], ph["tension_scale"], ph["adhesion_scale"], ph["xsph_rate"])
    o.set_state(sc.pos4, C["v0"] if v0 is None else v0)
    dual = precision in (96, 128)
    if dv is not None:
        t = np.zeros((sc.n, 4))
        t[:, :3] = dv
        o.set_tangent("vel", t)
    if dmu:
        o.seed_mu(dmu)
    blocks = fad.steps_for(grav, C["dt"], C["vc"])
    for k in range(len(grav)):
        if dual:
            o.seed_force(np.zeros(3) if dg is None else dg[k])
        o.substep(blocks[k], C["refine"])
    x = o.state_f64()["pos"][:, :3]
    if not dual:
        o.close()
        return _loss_of(x, C), None
    xt = torch.tensor(x, requires_grad=True)
    L = fad.loss(xt, sc.half, C["c_t"], torch.tensor(C["h_t"]), xp=torch)
    L.backward()
    d = float((xt.grad.numpy() * o.tangent("pos")[:, :3]).sum())
    o.close()
    return float(L.detach()), d


def _close(got, want, rel, what):
    _say(f"{what}: {got:.6e} against {want:.6e} ({abs(got - want) / abs(want):.2e} relative, bar {rel})")
    assert abs(got - want) <= rel * abs(want), what


@pytest.mark.timeout(900)
@pytest.mark.parametrize("physics", ["full", "viscous"])
def test_adjoint_64k_is_the_exact_derivative_of_the_fp32_step(physics):
    """The bench's adjoint workload: 64 substeps of the reference schedule
    at this spacing (4.17 ms each, R = 5), a checkpoint every 8 (eight taped
    segment replays). The gradient's directional derivatives along seeded
    directions of the initial velocities, the viscosity and the 64 gravity
    vectors equal the forward-mode derivative of the same FP32 trajectory
    (the oracle's float-valued dual numbers) to 1e-3.

    At this substep the step's linearisation is unstable: a tangent grows
    about e^0.33 a substep past the tenth, to ~1e12 here, so derivatives of
    the FP64 trajectory, or central differences, describe other trajectories
    and are printed, not held to the adjoint (the next test holds them at a
    substep where the linearisation is stable)."""
    steps, every = 64, 8
    C = _adjoint_config(steps=steps, physics=physics)
    sc = C["sc"]
    _say(f"adjoint_64k ({physics}): {sc.n} particles, {steps} substeps of {C['dt'] * 1e3:.2f} ms at R = "
         f"{C['refine']}, checkpoint every {every}")
    G = _gpu_gradient(C, every)
    rng = np.random.default_rng(7)
    u = rng.normal(0.0, 1.0, (sc.n, 3))
    ug = rng.normal(0.0, 1.0, (steps, 3))
    _close(float((G["v0"] * u).sum()), _oracle_run(C, 96, dv=u)[1], 1e-3, "v0 . u against the fp32 tangent")
    _close(G["mu"], _oracle_run(C, 96, dmu=1.0)[1], 1e-3, "mu against the fp32 tangent")
    _close(float((G["g"] * ug).sum()), _oracle_run(C, 96, dg=ug)[1], 1e-3, "g . ug against the fp32 tangent")
    _say(f"(unstable linearisation) mu: fp64 tangent {_oracle_run(C, 128, dmu=1.0)[1]:.6e}")


@pytest.mark.timeout(900)
@pytest.mark.parametrize("physics", ["full", "viscous"])
def test_adjoint_64k_matches_fp64_and_central_differences_at_a_stable_substep(physics):
    """The same scene and loss over 64 substeps of 1.04 ms (1/120 s in 8, R =
    2, the reference's short-substep regime), a checkpoint every 8, where the
    step's linearisation stays stable: the adjoint equals the FP32 tangent to
    1e-3, and the FP64 step's derivative and central differences (eps = 1e-3;
    the FP64 oracle's step, and the GPU's FP32 step) to 5% — the FP32 and
    FP64 trajectories part by rounding over 64 substeps, which moves their
    derivatives by up to ~3% here. Directions: a smooth field of initial
    velocities, the viscosity, a seeded direction of all 64 gravity vectors,
    and the x gravity of substep 32. Along a per-particle random direction
    of the initial velocities, a 1e-3 step crosses the kinks of the pressure
    clamp and of the histogram, so there the adjoint is held to the FP32
    tangent alone."""
    steps, every = 64, 8
    C = _adjoint_config(frame_substeps=8, steps=steps, physics=physics)
    sc = C["sc"]
    _say(f"adjoint_64k ({physics}): {sc.n} particles, {steps} substeps of {C['dt'] * 1e3:.2f} ms at R = "
         f"{C['refine']}, checkpoint every {every}")
    G = _gpu_gradient(C, every)
    eps = 1e-3
    p = sc.pos4[:, :3].astype(np.float64)
    us = np.stack([np.ones(sc.n), 0.5 * np.ones(sc.n), (p[:, 0] - p[:, 0].mean()) / 0.3], 1)
    got = float((G["v0"] * us).sum())
    _close(got, _oracle_run(C, 96, dv=us)[1], 1e-3, "smooth v0 against the fp32 tangent")
    _close(got, _oracle_run(C, 128, dv=us)[1], 0.05, "smooth v0 against the fp64 tangent")
    uv = np.zeros_like(C["v0"])
    uv[:, :3] = us
    fd = (_gpu_loss(C, v0=C["v0"] + eps * uv) - _gpu_loss(C, v0=C["v0"] - eps * uv)) / (2 * eps)
    _close(got, fd, 0.05, "smooth v0 against fp32 central differences")
    u = np.random.default_rng(7).normal(0.0, 1.0, (sc.n, 3))
    _close(float((G["v0"] * u).sum()), _oracle_run(C, 96, dv=u)[1], 1e-3, "random v0 against the fp32 tangent")
    _close(G["mu"], _oracle_run(C, 96, dmu=1.0)[1], 1e-3, "mu against the fp32 tangent")
    _close(G["mu"], _oracle_run(C, 128, dmu=1.0)[1], 0.05, "mu against the fp64 tangent")
    fd = (_oracle_run(C, 64, mu=MU * (1 + eps))[0] - _oracle_run(C, 64, mu=MU * (1 - eps))[0]) / (2 * eps * MU)
    _close(G["mu"], fd, 0.05, "mu against fp64 central differences")
    ug = np.random.default_rng(8).normal(0.0, 1.0, (steps, 3))
    got = float((G["g"] * ug).sum())
    _close(got, _oracle_run(C, 96, dg=ug)[1], 1e-3, "g . ug against the fp32 tangent")
    _close(got, _oracle_run(C, 128, dg=ug)[1], 0.05, "g . ug against the fp64 tangent")
    k = steps // 2
    dg = np.zeros_like(C["grav"])
    dg[k, 0] = eps * 9.81
    fd64 = (_oracle_run(C, 64, grav=C["grav"] + dg)[0] - _oracle_run(C, 64, grav=C["grav"] - dg)[0]) / (2 * eps * 9.81)
    fd32 = (_gpu_loss(C, grav=C["grav"] + dg) - _gpu_loss(C, grav=C["grav"] - dg)) / (2 * eps * 9.81)
    _close(G["g"][k, 0], fd64, 0.05, f"g[{k}].x against fp64 central differences")
    _close(G["g"][k, 0], fd32, 0.05, f"g[{k}].x against fp32 central differences")
