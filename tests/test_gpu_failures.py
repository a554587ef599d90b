"""Failure paths of the step, forced from the host (the reference's in-band
guards, crates/fluid-core/src/sim_solve.wgsl and render.rs):

- a NaN velocity uploaded through the state upload spreads exactly as the
  oracle spreads it: every word equal, NaNs as NaNs (their sign and payload
  bits are the hardware's own); a NaN distance keeps its candidate as the
  shader's `r2 >= support_sq` skip does (:450);
- a NaN or infinite position spreads as the oracle spreads it, and the
  poisoned particles come to rest clamped on the walls (:735-739; the
  reference's own history, render.rs:3165-3168);
- the alpha denominator guard (:508): an isolated particle's alpha is
  rho / 1e-4 and its kappa zero;
- a body force far past the CFL clamp: every clamp counted as the oracle
  counts them (:726-729).
"""
import numpy as np
import pytest

import fluid_amd as fa
from oracle import OracleSim

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


def _same(a, b):
    """Equal words, NaN where the other is NaN."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def _both(sc, vel=None, pos=None, step=None, refine=5, substeps=1):
    pos = sc.pos4 if pos is None else pos
    with fa.Sim(sc.params) as s:
        s.upload_state(pos, vel)
        for _ in range(substeps):
            s.substeps(step, 1, refine)
        g = s.state()
        g["nbr_n"] = s.grid()["nbr_n"]
        g["counters"] = s.counters()
        g["alpha"], g["density"], g["kd"] = s.field("alpha"), s.field("density"), s.field("kd")
    o = OracleSim(sc.params)
    o.set_state(pos, vel)
    for _ in range(substeps):
        o.substep(step, refine)
    r = o.state()
    r["nbr_n"] = o.grid()["nbr_n"]
    r["counters"] = o.counters()
    o.close()
    return g, r


def _step(sc, force=None):
    n, dt, refine, vc = sc.schedule()
    st = fa.pack_step(force=force or (0.0, -9.81, 0.0), dt=dt, v_clamp=vc)
    return st, refine


def test_a_nan_velocity_spreads_as_the_oracle_spreads_it():
    sc = fa.scenes.dam_break(seed=7)
    vel = np.zeros((sc.n, 4), np.float32)
    vel[sc.n // 2, 0] = np.nan
    st, refine = _step(sc)
    g, r = _both(sc, vel=vel, step=st, refine=refine, substeps=2)
    for k in STATE:
        assert _same(g[k], r[k]), k
    bad = np.isnan(g["vel"]).any(1)
    print(f"NaN velocities after 2 substeps: {int(bad.sum())} particles")
    assert 1 < bad.sum() < sc.n  # it spreads to the neighbours, and only there
    assert np.array_equal(g["nbr_n"], r["nbr_n"]) and g["counters"] == r["counters"]


@pytest.mark.parametrize("value", [np.nan, np.inf, -np.inf])
def test_a_non_finite_position_spreads_as_the_oracle_spreads_it(value):
    """A NaN or infinite coordinate sorts into the grid's guard layer (cell
    coordinate 0 or dims - 1 on that axis, sim_grid.wgsl:36-42) and every
    particle whose stencil holds that cell keeps it as a neighbour (the
    distance is unordered or infinite and the shader skips only r2 >=
    support_sq, sim_solve.wgsl:450): the same words as the oracle, NaNs as
    NaNs. Integrate's wall clamp then takes every non-finite coordinate to a
    wall (max(NaN, lo) is lo), so the resting positions are finite again and
    the poisoned particles sit on the walls."""
    sc = fa.scenes.dam_break(seed=8)
    pos = sc.pos4.copy()
    pos[sc.n // 3, 1] = value
    st, refine = _step(sc)
    g, r = _both(sc, pos=pos, step=st, refine=refine)
    for k in STATE:
        assert _same(g[k], r[k]), k
    assert np.array_equal(g["nbr_n"], r["nbr_n"]) and g["counters"] == r["counters"]
    assert np.isfinite(g["pos"][:, :3]).all()
    bad = np.zeros(sc.n, bool)
    for k in STATE:
        bad |= np.isnan(g[k].reshape(sc.n, -1)).any(1)
    print(f"position {value}: {int(bad.sum())} particles carry NaN after one substep")
    assert bad.any()


def test_the_alpha_guard_holds_an_isolated_particle():
    sc = fa.scenes.small_block((1, 1, 1), seed=1, half=(0.2, 0.2, 0.2))
    sc.pos4[0, :3] = (0.02, 0.01, -0.03)  # far from every wall: no fill, no wall gradient
    st, refine = _step(sc)
    g, r = _both(sc, step=st, refine=refine)
    for k in STATE:
        assert _same(g[k], r[k]), k
    assert g["alpha"][0] == np.float32(g["density"][0] / np.float32(1e-4))
    assert g["kd"][0] == 0.0


def test_a_violent_body_force_is_clamped_and_counted_like_the_oracle():
    sc = fa.scenes.dam_break(seed=9)
    st, refine = _step(sc, force=(400.0, -900.0, 50.0))
    g, r = _both(sc, step=st, refine=refine, substeps=3)
    for k in STATE:
        assert _same(g[k], r[k]), k
    assert g["counters"] == r["counters"] and g["counters"][0] > sc.n
    n, dt, _, vc = sc.schedule()
    sp = np.sqrt((g["vel"][:, :3].astype(np.float64) ** 2).sum(1))
    assert sp.max() <= 3 * vc * (1 + 1e-6)  # at most the spray ceiling
