"""Tolerance mode through the C ABI: DFSPH's divergence and density solves
iterated to their tolerances (the reference fixes one divergence pass and
`refine` density passes, sim_solve.wgsl:508-516, 665-703, render.rs:77-79;
the solves it iterates are those same kernels).

- a single sim in tolerance mode gives the oracle's tolerance mode every
  state word and the same pass counts, substep by substep, for the density
  solve alone, the divergence solve alone and both;
- a batch gives each scene the passes and bits a lone sim gives it, eager
  and under graph replay;
- the pass histograms count what the substeps ran;
- with the configs' tolerances (0.1% divergence, 0.01% density) every scene
  of a developed ensemble converges below the 256-pass cap.
"""
import numpy as np
import pytest

import fluid_amd as fa
from oracle import OracleSim

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _plan(sc, k):
    n, dt, r, vc = sc.schedule()
    return sc.step_block(k * dt, dt, vc, seed=k), r


SETTINGS = {
    "density": (1e-4, 1, 64, 0.0, 0),
    "divergence": (0.0, 1, 0, 1e-3, 64),
    "both": (1e-4, 2, 64, 1e-3, 64),
    "tight_caps": (1e-7, 1, 6, 1e-6, 3),
}


@pytest.mark.parametrize("mode", list(SETTINGS))
def test_single_sim_tolerance_is_bit_exact_with_the_oracle(mode):
    sc = fa.scenes.dam_break(seed=41)
    setting = SETTINGS[mode]
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    o.set_tolerance(*setting)
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        s.set_tolerance(*setting)
        s.pass_histogram(reset=True)
        seen = []
        for k in range(8):
            st, r = _plan(sc, k)
            s.substeps(st, 1, r)
            o.substep(st, r)
            got, (want, hist) = s.iterations(), o.passes()
            assert got == want, (k, got, want)
            seen.append(got)
            g, w = s.state(), o.state()
            for key in STATE:
                assert np.array_equal(_bits(g[key]), _bits(w[key])), (mode, k, key)
        assert np.array_equal(s.pass_histogram(), hist.astype(np.uint32))
        assert s.counters() == o.counters()
    o.close()
    print(f"{mode}: (density, divergence) passes by substep {seen}")
    den, div = zip(*seen)
    tol, lo, hi, dtol, dhi = setting
    if tol > 0:
        assert all(lo <= x <= hi for x in den)
    else:
        assert all(x == _plan(sc, 0)[1] for x in den)
    if dtol > 0:
        assert all(1 <= x <= dhi for x in div) and max(div) > 1
    else:
        assert all(x == 1 for x in div)


def test_tolerance_off_is_the_reference_step():
    sc = fa.scenes.dam_break(seed=42)
    st, r = _plan(sc, 0)
    out = []
    for setting in [None, (0.0, 1, 0, 0.0, 0)]:
        with fa.Sim(sc.params) as s:
            s.upload_state(sc.pos4)
            if setting:
                s.set_tolerance(1e-4, 1, 64, 1e-3, 64)
                s.set_tolerance(*setting)
            s.substeps(st, 2, r)
            out.append(s.state())
    for key in STATE:
        assert np.array_equal(_bits(out[0][key]), _bits(out[1][key])), key


def test_bad_tolerance_settings_are_refused():
    sc = fa.scenes.small_block((4, 4, 4), seed=43)
    with fa.Sim(sc.params) as s:
        for bad in [(1e-4, 3, 2, 0.0, 0), (1e-4, 1, 257, 0.0, 0), (1e-4, 1, 0, 0.0, 0), (0.0, 1, 0, 1e-3, 0),
                    (0.0, 1, 0, 1e-3, 257)]:
            with pytest.raises(RuntimeError):
                s.set_tolerance(*bad)
    with fa.Batch([sc.params]) as b:
        with pytest.raises(RuntimeError):
            b.set_tolerance(1e-4, 1, 257)
        with pytest.raises(RuntimeError):
            b.set_div_tolerance(1e-3, 257)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_tolerance_matches_lone_sims(graph):
    scenes = [
        fa.scenes.small_block((9, 7, 11), seed=44),
        fa.scenes.fill_block("fb_a", 3001, 0.3, seed=45),
        fa.scenes.small_block((1, 1, 1), seed=46),
        fa.scenes.dam_break(seed=47),
        fa.scenes.fill_block("fb_b", 2048, 0.45, seed=48, h_factor=1.2),
    ]
    rounds = 4
    plans = [[_plan(sc, k) for sc in scenes] for k in range(rounds)]
    with fa.Batch([sc.params for sc in scenes]) as b:
        b.use_graph(graph)
        b.set_tolerance(1e-4, 1, 64)
        b.set_div_tolerance(1e-3, 64)
        b.pass_histogram(reset=True)
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        its = []
        for k in range(rounds):
            b.substep([p[0] for p in plans[k]], [p[1] for p in plans[k]])
            its.append((b.iterations().copy(), b.div_iterations().copy()))
        states = [b.state(i) for i in range(len(scenes))]
        hist = b.pass_histogram()
    want = np.zeros((2, fa.HIST_BINS), np.int64)
    for i, sc in enumerate(scenes):
        with fa.Sim(sc.params) as s:
            s.upload_state(sc.pos4)
            s.set_tolerance(1e-4, 1, 64, 1e-3, 64)
            for k in range(rounds):
                s.substeps(plans[k][i][0], 1, plans[k][i][1])
                den, div = s.iterations()
                assert (den, div) == (its[k][0][i], its[k][1][i]), (sc.name, k)
                want[0, den] += 1
                want[1, div] += 1
            st = s.state()
        for key in STATE:
            assert np.array_equal(_bits(st[key]), _bits(states[i][key])), (sc.name, key)
    assert np.array_equal(hist, want)
    print("density passes by substep and scene:", [list(map(int, x[0])) for x in its])
    print("divergence passes by substep and scene:", [list(map(int, x[1])) for x in its])


def test_persistent_passes_give_the_words_of_a_launch_a_pass():
    """A batch's tolerance mode runs each solve as one persistent cooperative
    launch by default; a launch a pass gives the same passes and words."""
    scenes = [fa.scenes.fill_block(f"fb{i}", n, f, seed=60 + i)
              for i, (n, f) in enumerate([(2100, 0.3), (9000, 0.45), (30000, 0.25), (4096, 0.5)])]
    plans = [[_plan(sc, k) for sc in scenes] for k in range(3)]
    out = []
    for persistent in (True, False):
        with fa.Batch([sc.params for sc in scenes]) as b:
            b.persistent_passes(persistent)
            b.set_tolerance(1e-4, 1, 64)
            b.set_div_tolerance(1e-3, 32)
            for i, sc in enumerate(scenes):
                b.upload_state(i, sc.pos4)
            its = []
            for k in range(3):
                b.substep([p[0] for p in plans[k]], [p[1] for p in plans[k]])
                its.append((b.iterations().copy(), b.div_iterations().copy()))
            out.append((its, [b.state(i) for i in range(len(scenes))], b.pass_histogram()))
    (its_a, st_a, h_a), (its_b, st_b, h_b) = out
    print("persistent passes:", [list(map(int, x[0])) for x in its_a], [list(map(int, x[1])) for x in its_a])
    for a, b_ in zip(its_a, its_b):
        assert np.array_equal(a[0], b_[0]) and np.array_equal(a[1], b_[1])
    assert np.array_equal(h_a, h_b)
    for i in range(len(scenes)):
        for key in STATE:
            assert np.array_equal(_bits(st_a[i][key]), _bits(st_b[i][key])), (i, key)


def test_configs_tolerances_converge_below_the_cap_in_a_developed_ensemble():
    """The configs' tolerances, 0.1% divergence and 0.01% density, on 32
    ensemble scenes after 30 reference substeps: every scene and substep
    converges below the 256-pass cap, and a sample of scenes keeps the bits
    of the oracle in the same mode."""
    ens = fa.scenes.ensemble(32, seed=3)
    sched = [sc.schedule() for sc in ens]
    with fa.Batch([sc.params for sc in ens]) as b:
        b.use_graph(True)
        for i, sc in enumerate(ens):
            b.upload_state(i, sc.pos4)
        for k in range(30):
            b.substep([sc.step_block(k * sd[1], sd[1], sd[3], seed=k) for sc, sd in zip(ens, sched)],
                      [sd[2] for sd in sched])
        start = [b.state(i) for i in range(len(ens))]
        b.set_tolerance(1e-4, 1, 256)
        b.set_div_tolerance(1e-3, 256)
        b.pass_histogram(reset=True)
        steps = []
        for k in range(30, 36):
            steps.append([sc.step_block(k * sd[1], sd[1], sd[3], seed=k) for sc, sd in zip(ens, sched)])
            b.substep(steps[-1], [sd[2] for sd in sched])
        hist = b.pass_histogram()
        end = [b.state(i) for i in range(len(ens))]
    den, div = hist
    print("density passes histogram:", {p: int(c) for p, c in enumerate(den) if c})
    print("divergence passes histogram:", {p: int(c) for p, c in enumerate(div) if c})
    assert den.sum() == div.sum() == 6 * len(ens)
    assert den[256] == 0 and div[256] == 0
    order = np.argsort([sc.n for sc in ens])
    for i in [int(order[0]), int(order[-1]), int(order[len(order) // 2])]:
        sc = ens[i]
        o = OracleSim(sc.params)
        o.set_state(start[i]["pos"], start[i]["vel"], start[i]["prev_vel"], start[i]["prev_pressure"],
                    start[i]["temperature"])
        o.set_tolerance(1e-4, 1, 256, 1e-3, 256)
        for st in steps:
            o.substep(st[i], sched[i][2])
        w = o.state()
        o.close()
        for key in STATE:
            assert np.array_equal(_bits(end[i][key]), _bits(w[key])), (i, sc.n, key)
