"""Pins the oracle's tolerance mode (the divergence and density solves
iterated to a tolerance; the reference fixes one divergence pass and
`refine` density passes, sim_solve.wgsl:508-516, 665-703) by float64 brute
force: the pass count it picks is the first pass whose mean error, computed
here from the pair sums of Bender & Koschier (2015, alg. 2 and 3), is at or
under the tolerance, and the passes it runs are the reference schedule's
kernels (a tolerance run of P density passes is bit for bit the fixed
schedule with refine P).
"""
import numpy as np

import fluid_amd as fa
from oracle import OracleSim
from test_oracle_pinning_step import _consts, _kappa_scene, _pairs, _step

DT = float(np.float32(1.0 / 240.0))
ST = _step(force=(0.0, -9.81, 0.0), dt=DT, v_clamp=1e3)
HEAD = ("grid", "density_div", "div_apply", "forces_eval", "forces_den_apply")


def _drho(C, gradW, v, wg):
    """D rho / Dt = sum_j m (v_i - v_j) . grad W_ij + v_i . wall_grad."""
    return C["m"] * np.einsum("ijk,ijk->i", v[:, None, :] - v[None, :, :], gradW) + (v * wg).sum(1)


def _fresh(sc, vel):
    o = OracleSim(sc.params, precision=32)
    o.set_state(sc.pos4, vel)
    return o


def test_density_passes_stop_at_the_first_error_under_tol():
    sc, vel = _kappa_scene()
    C = _consts(sc)
    o = _fresh(sc, vel)
    for k in HEAD:
        o.kernel(k, ST)
    pos = o.field("work_pos")[:, :3]
    _, _, _, _, gradW, _ = _pairs(C, pos)
    rho, wg = o.field("density"), o.field("wall_grad")[:, :3]
    errors = []
    for _ in range(8):
        v = o.field("work_vel")[:, :3]
        excess = np.maximum(rho + DT * _drho(C, gradW, v, wg) - C["rho0"], 0.0)
        errors.append(excess.mean() / C["rho0"])
        o.kernel("den_kappa", ST)
        o.kernel("den_apply", ST)
    o.close()
    for k in (0, 2, 4):
        tol = float(np.sqrt(errors[k] * errors[k + 1]))  # far from both errors
        for lo, want in ((1, k + 2), (k + 3, k + 3)):
            t = _fresh(sc, vel)
            t.set_tolerance(tol, lo, 64)
            t.substep(ST, 5)
            (den, div), hist = t.passes()
            assert (den, div) == (want, 1), (k, lo, den, errors)
            assert hist[0, want] == 1 and hist[1, 1] == 1 and hist.sum() == 2
            f = _fresh(sc, vel)
            f.substep(ST, want)
            a, b = t.state(), f.state()
            for key in a:
                assert np.array_equal(a[key].view(np.uint32), b[key].view(np.uint32)), (k, key)
            t.close()
            f.close()


def test_divergence_passes_stop_at_the_first_error_under_tol():
    """Divergence pass r measures max(D rho / Dt, 0) dt / rho0 of the
    velocities r div_applies left and sets kd = max(D rho / Dt, 0) alpha /
    dt / rho (density_div's formula, sim_solve.wgsl:508-516)."""
    sc, vel = _kappa_scene()
    C = _consts(sc)
    o = _fresh(sc, vel)
    o.kernel("grid", ST)
    o.kernel("density_div", ST)
    pos = o.field("work_pos")[:, :3]
    _, _, _, _, gradW, _ = _pairs(C, pos)
    rho, alpha, wg = o.field("density"), o.field("alpha"), o.field("wall_grad")[:, :3]
    errors = []
    for r in range(8):
        v = o.field("work_vel")[:, :3]
        d = np.maximum(_drho(C, gradW, v, wg), 0.0)
        errors.append((d * DT / C["rho0"]).mean())
        if r:
            o.kernel("div_kappa", ST)
        want = d * alpha / DT / rho
        err = np.linalg.norm(o.field("kd") - want) / np.linalg.norm(want)
        assert r >= 3 or err < 1e-5, (r, err)
        o.kernel("div_apply", ST)
    o.close()
    assert errors[0] > 1e-3 and errors[3] < errors[0]
    for k in (0, 1, 3):
        tol = float(np.sqrt(errors[k] * errors[k + 1]))
        t = _fresh(sc, vel)
        t.set_tolerance(0.0, 1, 0, tol, 64)
        t.substep(ST, 5)
        (den, div), _ = t.passes()
        assert (den, div) == (5, k + 2), (k, den, div, errors)
        t.close()
    # the cap holds whatever the error
    t = _fresh(sc, vel)
    t.set_tolerance(0.0, 1, 0, 1e-30, 3)
    t.substep(ST, 5)
    assert t.passes()[0] == (5, 3)
    t.close()


def test_tolerance_off_is_the_reference_substep():
    sc = fa.scenes.dam_break(seed=5)
    n, dt, r, vc = sc.schedule()
    st = fa.pack_step(force=(0.0, -9.81, 0.0), dt=dt, v_clamp=vc)
    a, b = OracleSim(sc.params), OracleSim(sc.params)
    for o in (a, b):
        o.set_state(sc.pos4)
    b.set_tolerance(1e-4, 1, 64, 1e-3, 64)
    b.set_tolerance(0.0, 1, 0, 0.0, 0)
    a.substep(st, r)
    b.substep(st, r)
    assert b.passes()[0] == (r, 1) and b.passes()[1].sum() == 0
    sa, sb = a.state(), b.state()
    for key in sa:
        assert np.array_equal(sa[key].view(np.uint32), sb[key].view(np.uint32)), key
    a.close()
    b.close()
