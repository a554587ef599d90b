"""The adjoint of the step (torch.autograd.Function over the C ABI) against
the FP64 reference gradient and central finite differences.

The FP64 reference gradient is the dual-number run of the oracle: the
oracle's double-precision step evaluated on (value, derivative) pairs, which
gives the exact directional derivative of the FP64 step along a seeded
direction of the initial velocities, the viscosity or the gravity vectors.
The GPU adjoint differentiates the FP32 step; its directional derivatives
are held to 2% of the FP64 reference (the two agree to about 1e-6 here; the
margin covers FP32 drift on longer runs) and to central differences with
eps = 1e-3: of the FP64 step for the viscosity and the gravity, of the GPU's
FP32 step for the gravity and the initial velocities. (The FP32 loss rounds
at ~3e-5 of ~400, above the effect of a 1e-3 change of mu, so mu is
differenced in FP64.)
"""
import numpy as np
import pytest

import fluid_amd as fa
from fluid_amd import adjoint as fad
from oracle import OracleSim

pytestmark = pytest.mark.gpu

MU = 1.002e-3


def _setup(steps=8, n=2000, seed=41):
    sc = fa.scenes.fill_block("adj", n, 0.3, seed=seed)
    nsub, dt, refine, vc = sc.schedule()
    rng = np.random.default_rng(seed)
    v0 = np.zeros((sc.n, 4), np.float32)
    v0[:, :3] = rng.normal(0.0, 0.1, (sc.n, 3))
    th = rng.normal(0.0, 0.2, steps)
    grav = np.stack([9.81 * np.sin(th), -9.81 * np.cos(th), np.zeros(steps)], 1)
    half = sc.half
    c_t = sc.pos4[:, :3].mean(0) + np.array([0.02, 0.01, -0.01])
    shifted = sc.pos4[:, :3].astype(np.float64) + np.array([0.03, 0.0, 0.0])
    h_t = fad.soft_histogram(shifted, half)
    return sc, dt, refine, vc, v0, grav, half, c_t, h_t


def _gpu_loss_and_grad(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, every, mu=MU):
    import torch
    adj = fad.Adjoint(sc.params, steps, every)
    Sim = fad.simulate_function()
    v = torch.tensor(v0[:, :3], dtype=torch.float32, requires_grad=True)
    m = torch.tensor(mu, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(grav, dtype=torch.float64, requires_grad=True)
    x = Sim.apply(v, m, g, adj, sc.pos4, dt, vc, refine).double()
    L = fad.loss(x, half, c_t, torch.tensor(h_t), xp=torch)
    L.backward()
    fwd_ms, bwd_ms = adj.timing()
    adj.close()
    return float(L), v.grad.numpy().astype(np.float64), float(m.grad), g.grad.numpy(), fwd_ms, bwd_ms


def _gpu_loss(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, mu=MU):
    adj = fad.Adjoint(sc.params, steps, steps)
    adj.set_mu(mu)
    pos, _ = adj.forward(sc.pos4, v0, fad.steps_for(grav, dt, vc), [refine] * steps)
    adj.close()
    return float(fad.loss(pos[:, :3].astype(np.float64), half, c_t, h_t))


def _dual_derivative(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, dv=None, dmu=0.0, dg=None, mu=MU):
    """d loss along the seeded direction, through the FP64 dual oracle."""
    import torch
    o = OracleSim(sc.params, precision=128)
    o.set_physics(mu, 0.0, 0.0, 0.0, 0.0)
    o.set_state(sc.pos4, v0)
    if dv is not None:
        t = np.zeros((sc.n, 4))
        t[:, :3] = dv
        o.set_tangent("vel", t)
    if dmu:
        o.seed_mu(dmu)
    blocks = fad.steps_for(grav, dt, vc)
    for k in range(steps):
        o.seed_force(np.zeros(3) if dg is None else dg[k])
        o.substep(blocks[k], refine)
    x = torch.tensor(o.state_f64()["pos"][:, :3], requires_grad=True)
    L = fad.loss(x, half, c_t, torch.tensor(h_t), xp=torch)
    L.backward()
    return float((x.grad.numpy() * o.tangent("pos")[:, :3]).sum())


@pytest.mark.parametrize("steps,every", [(1, 1), (6, 2)])
def test_adjoint_matches_fp64_reference_gradient(steps, every):
    sc, dt, refine, vc, v0, grav, half, c_t, h_t = _setup(steps)
    L, g_v0, g_mu, g_g, fwd_ms, bwd_ms = _gpu_loss_and_grad(sc, dt, refine, vc, v0, grav[:steps], half, c_t, h_t,
                                                           steps, every)
    rng = np.random.default_rng(7)
    u = rng.normal(0.0, 1.0, (sc.n, 3))
    ug = rng.normal(0.0, 1.0, (steps, 3))
    ref_v = _dual_derivative(sc, dt, refine, vc, v0, grav[:steps], half, c_t, h_t, steps, dv=u)
    ref_mu = _dual_derivative(sc, dt, refine, vc, v0, grav[:steps], half, c_t, h_t, steps, dmu=1.0)
    ref_g = _dual_derivative(sc, dt, refine, vc, v0, grav[:steps], half, c_t, h_t, steps, dg=ug)
    got_v, got_g = float((g_v0 * u).sum()), float((g_g * ug).sum())
    print(f"loss {L:.6e}  v0: adjoint {got_v:.6e} fp64 {ref_v:.6e}  mu: adjoint {g_mu:.6e} fp64 {ref_mu:.6e}  "
          f"g: adjoint {got_g:.6e} fp64 {ref_g:.6e}  forward {fwd_ms:.2f} ms backward {bwd_ms:.2f} ms")
    assert abs(got_v - ref_v) <= 0.02 * abs(ref_v) + 1e-9
    assert abs(g_mu - ref_mu) <= 0.02 * abs(ref_mu) + 1e-9
    assert abs(got_g - ref_g) <= 0.02 * abs(ref_g) + 1e-9


def _fp64_loss(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, mu=MU):
    """The loss after the FP64 oracle's run of the same substeps."""
    o = OracleSim(sc.params, precision=64)
    o.set_physics(mu, 0.0, 0.0, 0.0, 0.0)
    o.set_state(sc.pos4, v0)
    blocks = fad.steps_for(grav, dt, vc)
    for k in range(steps):
        o.substep(blocks[k], refine)
    return float(fad.loss(o.state_f64()["pos"][:, :3], half, c_t, h_t))


def test_adjoint_matches_central_differences():
    """Central differences (eps = 1e-3) of the FP64 step for the viscosity
    (relative) and one gravity component (relative to 9.81), and of the GPU
    FP32 step for the gravity and the initial velocities, against the GPU
    adjoint. The max(., 0) clamps of the pressure solve make the loss
    piecewise smooth; 5% covers the kinks a difference step can cross."""
    steps = 4
    sc, dt, refine, vc, v0, grav, half, c_t, h_t = _setup(steps)
    grav = grav[:steps]
    L, g_v0, g_mu, g_g, _, _ = _gpu_loss_and_grad(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, 2)
    eps = 1e-3
    lp = _fp64_loss(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, mu=MU * (1 + eps))
    lm = _fp64_loss(sc, dt, refine, vc, v0, grav, half, c_t, h_t, steps, mu=MU * (1 - eps))
    fd_mu = (lp - lm) / (2 * eps * MU)
    dg = np.zeros_like(grav)
    dg[1, 0] = eps * 9.81
    lp = _fp64_loss(sc, dt, refine, vc, v0, grav + dg, half, c_t, h_t, steps)
    lm = _fp64_loss(sc, dt, refine, vc, v0, grav - dg, half, c_t, h_t, steps)
    fd_g64 = (lp - lm) / (2 * eps * 9.81)
    lp = _gpu_loss(sc, dt, refine, vc, v0, grav + dg, half, c_t, h_t, steps)
    lm = _gpu_loss(sc, dt, refine, vc, v0, grav - dg, half, c_t, h_t, steps)
    fd_g32 = (lp - lm) / (2 * eps * 9.81)
    u = np.zeros_like(v0)
    u[:, :3] = np.random.default_rng(9).normal(0.0, 1.0, (sc.n, 3))
    lp = _gpu_loss(sc, dt, refine, vc, v0 + eps * u, grav, half, c_t, h_t, steps)
    lm = _gpu_loss(sc, dt, refine, vc, v0 - eps * u, grav, half, c_t, h_t, steps)
    fd_v32 = (lp - lm) / (2 * eps)
    got_v = float((g_v0 * u[:, :3]).sum())
    print(f"mu: adjoint {g_mu:.6e} fd64 {fd_mu:.6e}   g[1].x: adjoint {g_g[1, 0]:.6e} fd64 {fd_g64:.6e} "
          f"fd32 {fd_g32:.6e}   v0: adjoint {got_v:.6e} fd32 {fd_v32:.6e}")
    assert abs(g_mu - fd_mu) <= 0.05 * abs(fd_mu) + 1e-9
    assert abs(g_g[1, 0] - fd_g64) <= 0.05 * abs(fd_g64) + 1e-9
    assert abs(g_g[1, 0] - fd_g32) <= 0.05 * abs(fd_g32) + 1e-9
    assert abs(got_v - fd_v32) <= 0.10 * abs(fd_v32) + 1e-9


def test_backward_after_a_second_forward_on_the_same_adjoint_raises():
    """The device tapes belong to an Adjoint's last forward: the output of an
    earlier forward on it refuses a backward instead of returning the later
    run's gradient; the later output still differentiates."""
    import torch
    steps = 2
    sc, dt, refine, vc, v0, grav, half, c_t, h_t = _setup(steps, n=600)
    adj = fad.Adjoint(sc.params, steps, 1)
    Sim = fad.simulate_function()
    v = torch.tensor(v0[:, :3], dtype=torch.float32, requires_grad=True)
    m = torch.tensor(MU, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(grav[:steps], dtype=torch.float64, requires_grad=True)
    first = Sim.apply(v, m, g, adj, sc.pos4, dt, vc, refine)
    second = Sim.apply(v, m, g * 1.01, adj, sc.pos4, dt, vc, refine)
    with pytest.raises(RuntimeError, match="another forward"):
        first.sum().backward()
    second.sum().backward()
    assert np.isfinite(v.grad.numpy()).all() and np.abs(v.grad.numpy()).sum() > 0
    adj.close()


# The full step's terms one at a time over the viscous profile, then all of
# them in a turning frame: near pressure, cohesion, wall adhesion, the XSPH
# blend, the rotating frame's fictitious forces.
FRAME = ((0.0, 0.0, 0.8), (0.0, 0.0, 0.3))
NO_FRAME = ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
PROFILES = {
    "near": (dict(fad.VISCOUS, k_near=3000.0), NO_FRAME),
    "cohesion": (dict(fad.VISCOUS, tension_scale=1.0), NO_FRAME),
    "adhesion": (dict(fad.VISCOUS, adhesion_scale=1.0), NO_FRAME),
    "xsph": (dict(fad.VISCOUS, xsph_rate=6.0), NO_FRAME),
    "rotation": (fad.VISCOUS, FRAME),
    "full": (fad.FULL, FRAME),
}


def _tangent(sc, dt, refine, vc, v0, grav, half, c_t, h_t, physics, frame, precision=96, dv=None, dmu=0.0,
             dg=None, mu=MU):
    """d loss along a seeded direction through the oracle's dual numbers:
    precision 96 is the FP32 step's own derivative (float values, double
    tangents: the trajectory the GPU runs, word for word), 128 the FP64
    step's."""
    import torch
    o = OracleSim(sc.params, precision=precision)
    o.set_physics(mu, physics["k_near"], physics["tension_scale"], physics["adhesion_scale"], physics["xsph_rate"])
    o.set_state(sc.pos4, v0)
    if dv is not None:
        t = np.zeros((sc.n, 4))
        t[:, :3] = dv
        o.set_tangent("vel", t)
    if dmu:
        o.seed_mu(dmu)
    blocks = fad.steps_for(grav, dt, vc, *frame)
    for k in range(len(grav)):
        o.seed_force(np.zeros(3) if dg is None else dg[k])
        o.substep(blocks[k], refine)
    x = torch.tensor(o.state_f64()["pos"][:, :3], requires_grad=True)
    L = fad.loss(x, half, c_t, torch.tensor(h_t), xp=torch)
    L.backward()
    d = float((x.grad.numpy() * o.tangent("pos")[:, :3]).sum())
    o.close()
    return d


@pytest.mark.parametrize("profile", list(PROFILES))
def test_full_step_adjoint_is_the_fp32_steps_derivative(profile):
    """The adjoint with each of the full step's terms switched on (and all of
    them in a turning frame) against the forward-mode derivative of the same
    FP32 trajectory, to 1e-3, along seeded directions of the initial
    velocities, the viscosity and the gravity vectors; the FP64 step's
    derivative to 2%. Each term moves the derivative along the initial
    velocities by far more than the bar (against the viscous profile's
    derivative), so a dropped or mis-signed term fails."""
    import torch
    physics, frame = PROFILES[profile]
    steps, every = 4, 2
    sc, dt, refine, vc, v0, grav, half, c_t, h_t = _setup(steps)
    grav = grav[:steps]
    adj = fad.Adjoint(sc.params, steps, every, physics=physics)
    Sim = fad.simulate_function()
    v = torch.tensor(v0[:, :3], dtype=torch.float32, requires_grad=True)
    m = torch.tensor(MU, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(grav, dtype=torch.float64, requires_grad=True)
    x = Sim.apply(v, m, g, adj, sc.pos4, dt, vc, refine, frame).double()
    L = fad.loss(x, half, c_t, torch.tensor(h_t), xp=torch)
    L.backward()
    adj.close()
    rng = np.random.default_rng(7)
    u = rng.normal(0.0, 1.0, (sc.n, 3))
    ug = rng.normal(0.0, 1.0, (steps, 3))
    case = (sc, dt, refine, vc, v0, grav, half, c_t, h_t, physics, frame)
    got = {"v0": float((v.grad.numpy().astype(np.float64) * u).sum()), "mu": float(m.grad),
           "g": float((g.grad.numpy() * ug).sum())}
    want = {"v0": _tangent(*case, dv=u), "mu": _tangent(*case, dmu=1.0), "g": _tangent(*case, dg=ug)}
    fp64 = {"v0": _tangent(*case, precision=128, dv=u)}
    base = _tangent(sc, dt, refine, vc, v0, grav, half, c_t, h_t, fad.VISCOUS, NO_FRAME, dv=u)
    print(f"{profile}: loss {float(L):.6e}; v0: adjoint {got['v0']:.6e} fp32 {want['v0']:.6e} "
          f"fp64 {fp64['v0']:.6e} viscous {base:.6e}; mu: adjoint {got['mu']:.6e} fp32 {want['mu']:.6e}; "
          f"g: adjoint {got['g']:.6e} fp32 {want['g']:.6e}")
    for key in got:
        assert abs(got[key] - want[key]) <= 1e-3 * abs(want[key]), key
    assert abs(got["v0"] - fp64["v0"]) <= 0.02 * abs(fp64["v0"])
    assert abs(want["v0"] - base) > 10 * 1e-3 * abs(want["v0"]), "the term does not move the derivative"


def test_adjoint_refuses_touches():
    sc = fa.scenes.fill_block("adj", 600, 0.3, seed=3)
    nsub, dt, refine, vc = sc.schedule()
    adj = fad.Adjoint(sc.params, 1, 1)
    adj.set_mu(MU)
    touched = fa.pack_step(force=(0.0, -9.81, 0.0), dt=dt, v_clamp=vc, touches=[(0.0, 0.0, 0.1, 0.0)], touch_n=1,
                           touch_r=0.2)
    with pytest.raises(RuntimeError, match="touches"):
        adj.forward(sc.pos4, None, [touched], [refine])
    adj.close()
