"""The batched step through the C ABI: a batch of uneven scenes gives each
scene the bits a lone Sim gives it (state, grid, fields, counters, stats),
in the reference schedule, under hipGraph replay, in tolerance mode and after
autotuning. A scene without particles, alone or in a batch, steps as a no-op.
"""
import numpy as np
import pytest

import fluid_amd as fa

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _scenes():
    # uneven counts (not multiples of 256), both smoothing-length readings,
    # and a lone particle
    return [
        fa.scenes.small_block((9, 7, 11), seed=31),
        fa.scenes.fill_block("fb_a", 3001, 0.3, seed=32),
        fa.scenes.small_block((1, 1, 1), seed=33),
        fa.scenes.dam_break(seed=34),
        fa.scenes.fill_block("fb_b", 2048, 0.45, seed=35, h_factor=1.2),
    ]


def _steps(scenes, k, dt_scale=None):
    out, refine = [], []
    for i, sc in enumerate(scenes):
        n, dt, r, vc = sc.schedule()
        if dt_scale is not None:
            dt *= dt_scale[i]
            r = fa.refine_passes(dt)
        out.append(sc.step_block(k * dt, dt, vc, seed=k))
        refine.append(r)
    return out, refine


def _lone(sc, steps):
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        for st, r in steps:
            s.substeps(st, 1, r)
        out = s.state()
        out["grid"] = s.grid()
        out["fields"] = {k: s.field(k) for k in ("density", "alpha", "kd", "pressure", "accel", "xsph")}
        out["counters"] = s.counters()
        s.reduce_stats()
        out["stats"] = s.read_stats()
        return out


def test_batch_matches_lone_sims_bit_for_bit():
    scenes = _scenes()
    scale = [1.0, 0.1, 1.0, 0.5, 1.0]  # mixes R = 5 and R = 2 in one launch
    plan = [_steps(scenes, k, scale) for k in range(3)]
    with fa.Batch([sc.params for sc in scenes]) as b:
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        for steps, refine in plan:
            b.substep(steps, refine)
        b.reduce_stats()
        stats = b.read_stats()
        counters = b.counters()
        for i, sc in enumerate(scenes):
            ref = _lone(sc, [(plan[k][0][i], plan[k][1][i]) for k in range(3)])
            st = b.state(i)
            for key in STATE:
                assert np.array_equal(_bits(st[key]), _bits(ref[key])), (sc.name, key)
            g = b.grid(i)
            for key in ("cellid", "counts", "starts", "perm"):
                assert np.array_equal(g[key], ref["grid"][key]), (sc.name, key)
            for key, v in ref["fields"].items():
                assert np.array_equal(_bits(b.field(i, key)), _bits(v)), (sc.name, key)
            assert tuple(int(x) for x in counters[i]) == ref["counters"], sc.name
            assert np.array_equal(_bits(stats[i]), _bits(ref["stats"])), sc.name


@pytest.mark.parametrize("graph,streams", [(False, 2), (True, 2), (False, 1), (True, 4)])
def test_scene_groups_give_the_bits_of_one_group(graph, streams, monkeypatch):
    """FLUID_BATCH_GROUP runs the sweeps group by group of whole scenes
    (here a group closes after 2,048 slots, so most scenes are a group of
    their own), in turn over FLUID_BATCH_STREAMS streams; scenes do not
    interact, so every word matches one group."""
    scenes = _scenes()
    scale = [1.0, 0.1, 1.0, 0.5, 1.0]
    plan = [_steps(scenes, k, scale) for k in range(3)]
    monkeypatch.setenv("FLUID_BATCH_STREAMS", str(streams))

    def run(group):
        monkeypatch.setenv("FLUID_BATCH_GROUP", str(group))
        with fa.Batch([sc.params for sc in scenes]) as b:
            b.use_graph(graph)
            for i, sc in enumerate(scenes):
                b.upload_state(i, sc.pos4)
            for steps, refine in plan:
                b.substep(steps, refine)
            return [(b.state(i), {k: b.field(i, k) for k in ("density", "kd", "pressure", "accel")})
                    for i in range(len(scenes))]

    one, grouped = run(0), run(2048)
    for i in range(len(scenes)):
        for key in STATE:
            assert np.array_equal(_bits(grouped[i][0][key]), _bits(one[i][0][key])), (i, key)
        for key in one[i][1]:
            assert np.array_equal(_bits(grouped[i][1][key]), _bits(one[i][1][key])), (i, key)


def test_graph_replays_match_eager_substeps():
    scenes = _scenes()[:4]
    plan = [_steps(scenes, k) for k in range(4)]

    def run(graph):
        states = []
        with fa.Batch([sc.params for sc in scenes]) as b:
            b.use_graph(graph)
            for i, sc in enumerate(scenes):
                b.upload_state(i, sc.pos4)
            for steps, refine in plan:
                b.substep(steps, refine)
                # eager work and a synchronisation between replays
                b.reduce_stats()
                states.append([b.state(i) for i in range(len(scenes))])
        return states

    eager, graph = run(False), run(True)
    for k in range(len(plan)):
        for i in range(len(scenes)):
            for key in STATE:
                assert np.array_equal(_bits(graph[k][i][key]), _bits(eager[k][i][key])), (k, i, key)


def test_scene_with_dt_zero_sits_out():
    scenes = _scenes()[:3]
    steps, refine = _steps(scenes, 0)
    steps[1] = fa.pack_step(force=(0.0, -9.81, 0.0), dt=0.0, v_clamp=1.0)
    with fa.Batch([sc.params for sc in scenes]) as b:
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        b.substep(steps, refine)
        assert np.array_equal(b.state(1)["pos"], scenes[1].pos4)
        ref = _lone(scenes[0], [(steps[0], refine[0])])
        assert np.array_equal(_bits(b.state(0)["pos"]), _bits(ref["pos"]))


def test_tolerance_mode_converges_per_scene_independently():
    scenes = _scenes()
    steps, _ = _steps(scenes, 0)
    with fa.Batch([sc.params for sc in scenes]) as b:
        b.set_tolerance(1e-4, 1, 12)
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        b.substep(steps, [0] * len(scenes))
        its = b.iterations()
        states = [b.state(i) for i in range(len(scenes))]
    assert all(1 <= x <= 12 for x in its), its
    # each scene's passes and bits are those of a batch of that scene alone
    for i, sc in enumerate(scenes):
        with fa.Batch([sc.params]) as one:
            one.set_tolerance(1e-4, 1, 12)
            one.upload_state(0, sc.pos4)
            one.substep([steps[i]], [0])
            assert one.iterations()[0] == its[i], sc.name
            st = one.state(0)
        for key in STATE:
            assert np.array_equal(_bits(st[key]), _bits(states[i][key])), (sc.name, key)
    # a loose tolerance stops every scene after min_passes
    with fa.Batch([sc.params for sc in scenes]) as b:
        b.set_tolerance(10.0, 2, 12)
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        b.substep(steps, [0] * len(scenes))
        assert list(b.iterations()) == [2] * len(scenes)


def test_tolerance_mode_with_fixed_passes_is_the_reference_schedule():
    """min_passes = max_passes = R runs exactly R passes: the same bits as the
    reference schedule with refine R."""
    scenes = _scenes()[:3]
    steps, _ = _steps(scenes, 0)
    with fa.Batch([sc.params for sc in scenes]) as b:
        b.set_tolerance(1e-9, 3, 3)
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        b.substep(steps, [0] * len(scenes))
        tol_states = [b.state(i) for i in range(len(scenes))]
    with fa.Batch([sc.params for sc in scenes]) as b:
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        b.substep(steps, [3] * len(scenes))
        for i in range(len(scenes)):
            st = b.state(i)
            for key in STATE:
                assert np.array_equal(_bits(st[key]), _bits(tol_states[i][key])), (i, key)


def test_autotune_restores_state_and_keeps_results():
    """The trial substeps leave no trace: the resting set, and the cumulative
    clamp and overflow counters, are those of lone sims that never ran them."""
    scenes = _scenes()
    plan = [_steps(scenes, k) for k in range(2)]
    with fa.Batch([sc.params for sc in scenes]) as b:
        for i, sc in enumerate(scenes):
            b.upload_state(i, sc.pos4)
        cap = b.autotune(plan[0][0], plan[0][1], trials=1)
        assert cap == b.describe()["stage_cap"]
        for i, sc in enumerate(scenes):
            assert np.array_equal(b.state(i)["pos"], sc.pos4)
        assert not b.counters().any()
        for steps, refine in plan:
            b.substep(steps, refine)
        tuned = [b.state(i) for i in range(len(scenes))]
        counters = b.counters()
    for i, sc in enumerate(scenes):
        ref = _lone(sc, [(plan[k][0][i], plan[k][1][i]) for k in range(2)])
        for key in STATE:
            assert np.array_equal(_bits(tuned[i][key]), _bits(ref[key])), (sc.name, key)
        assert tuple(int(x) for x in counters[i]) == ref["counters"], sc.name


def test_tolerance_changes_reach_graph_replays():
    """A captured substep holds the tolerance and the pass bounds as kernel
    arguments; changing them between graph substeps gives what an eager run
    with the same changes gives (passes and bits)."""
    scenes = _scenes()[:4]
    plan = [_steps(scenes, k)[0] for k in range(4)]
    settings = [(1e-2, 1, 12), (1e-6, 2, 12), (1e-2, 3, 12), (1e-6, 1, 5)]

    def run(graph):
        out = []
        with fa.Batch([sc.params for sc in scenes]) as b:
            b.use_graph(graph)
            for i, sc in enumerate(scenes):
                b.upload_state(i, sc.pos4)
            for steps, (tol, lo, hi) in zip(plan, settings):
                b.set_tolerance(tol, lo, hi)
                b.substep(steps, [0] * len(scenes))
                out.append((b.iterations().copy(), [b.state(i) for i in range(len(scenes))]))
        return out

    eager, graph = run(False), run(True)
    for k in range(len(plan)):
        assert np.array_equal(graph[k][0], eager[k][0]), (k, graph[k][0], eager[k][0])
        assert all(settings[k][1] <= x <= settings[k][2] for x in graph[k][0]), (k, graph[k][0])
        for i in range(len(scenes)):
            for key in STATE:
                assert np.array_equal(_bits(graph[k][1][i][key]), _bits(eager[k][1][i][key])), (k, i, key)


def _empty_params(sc):
    """sc's SimParams block with its particle count (bytes 28-31) set to 0."""
    p = bytearray(sc.params)
    p[28:32] = (0).to_bytes(4, "little")
    return bytes(p)


def test_an_empty_sim_steps_as_a_no_op():
    sc = fa.scenes.small_block((4, 4, 4), seed=41)
    (step,), (refine,) = _steps([sc], 0)
    with fa.Sim(_empty_params(sc)) as s:
        assert s.n == 0
        s.upload_state(np.zeros((0, 4), np.float32))
        s.substeps(step, 2, refine)
        s.density_only(step)
        s.frame(step, 1, refine)
        st = s.state()
        assert all(v.shape[0] == 0 for v in st.values())
        g = s.grid()
        assert not g["starts"].any()
        assert s.counters() == (0, 0)
        assert not s.read_stats().any()


def test_an_empty_scene_in_a_batch_leaves_the_others_bit_exact():
    scenes = [fa.scenes.small_block((5, 6, 7), seed=42), fa.scenes.small_block((4, 4, 4), seed=43),
              fa.scenes.fill_block("fb_c", 700, 0.3, seed=44)]
    params = [scenes[0].params, _empty_params(scenes[1]), scenes[2].params]
    plan = [_steps(scenes, k) for k in range(2)]
    with fa.Batch(params) as b:
        b.upload_state(0, scenes[0].pos4)
        b.upload_state(2, scenes[2].pos4)
        for steps, refine in plan:
            b.substep(steps, refine)
        b.reduce_stats()
        stats = b.read_stats()
        counters = b.counters()
        assert b.state(1)["pos"].shape[0] == 0
        assert not stats[1].any()
        assert tuple(int(x) for x in counters[1]) == (0, 0)
        for i in (0, 2):
            ref = _lone(scenes[i], [(plan[k][0][i], plan[k][1][i]) for k in range(2)])
            st = b.state(i)
            for key in STATE:
                assert np.array_equal(_bits(st[key]), _bits(ref[key])), (i, key)
            assert np.array_equal(b.grid(i)["perm"], ref["grid"]["perm"]), i
            assert np.array_equal(_bits(stats[i]), _bits(ref["stats"])), i
