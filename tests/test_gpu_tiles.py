"""The solve's tile size and the batch's launch shape are run-time choices
(three builds of the solve, 64, 128 and 256 slots a workgroup); none of
them changes a result bit:

- a sim gives the same words at every tile size, and the oracle's;
- a batch gives each scene the same words under every launch shape (tile,
  staging capacity, scene groups, streams), eager and under graph replay;
- the autotuner leaves a valid shape and the results of a batch that never
  ran its trials.
"""
import numpy as np
import pytest

import fluid_amd as fa
from oracle import OracleSim

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _run_sim(sc, tile, substeps=4, stage_bytes=None):
    with fa.Sim(sc.params) as s:
        s.set_tile(tile)
        assert s.tile == tile and s.describe()["tile"] == tile
        s.upload_state(sc.pos4)
        n, dt, r, vc = sc.schedule()
        for k in range(substeps):
            s.substeps(sc.step_block(k * dt, dt, vc, seed=k), 1, r)
        out = s.state()
        out["nbr"] = s.neighbours()
        out["counters"] = s.counters()
        return out


@pytest.mark.parametrize("maker", [lambda: fa.scenes.dam_break(seed=51),
                                   lambda: fa.scenes.fill_block("fb", 30001, 0.35, seed=52)])
def test_every_tile_size_gives_the_same_words_as_the_oracle(maker):
    sc = maker()
    runs = {t: _run_sim(sc, t) for t in (64, 128, 256)}
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    n, dt, r, vc = sc.schedule()
    for k in range(4):
        o.substep(sc.step_block(k * dt, dt, vc, seed=k), r)
    want = o.state()
    counters = o.counters()
    o.close()
    for t, got in runs.items():
        for key in STATE:
            assert np.array_equal(_bits(got[key]), _bits(want[key])), (t, key)
        assert np.array_equal(got["nbr"], runs[128]["nbr"]), t
        assert got["counters"] == counters, t


def test_a_bad_tile_size_is_refused():
    sc = fa.scenes.small_block((4, 4, 4), seed=53)
    with fa.Sim(sc.params) as s:
        with pytest.raises(RuntimeError):
            s.set_tile(96)
        assert s.tile == 128
    with fa.Batch([sc.params]) as b:
        with pytest.raises(RuntimeError):
            b.set_tuning(32, 1024, 0, 1)
        with pytest.raises(RuntimeError):
            b.set_tuning(128, 1024, 0, 5)


def _ensemble():
    return fa.scenes.ensemble(12, seed=7, n_min=1500, n_max=40000)


def _plan(ens, k):
    sched = [sc.schedule() for sc in ens]
    return ([sc.step_block(k * sd[1], sd[1], sd[3], seed=k) for sc, sd in zip(ens, sched)],
            [sd[2] for sd in sched])


SHAPES = [(64, 4096, 0, 1), (128, 0, 0, 1), (128, 1024, 65536, 2), (256, 4096, 100000, 3), (256, 1536, 40000, 4)]


@pytest.mark.parametrize("graph", [False, True])
def test_every_batch_launch_shape_gives_the_same_words(graph):
    ens = _ensemble()
    plans = [_plan(ens, k) for k in range(3)]
    results = []
    for shape in SHAPES:
        with fa.Batch([sc.params for sc in ens]) as b:
            b.set_tuning(*shape)
            t = b.tuning()
            assert t["tile"] == shape[0] and t["group_slots"] == shape[2] and t["streams"] == shape[3], t
            assert t["stage_cap"] <= shape[1] and t["stage_cap"] % 64 == 0, t
            d = b.describe()
            assert d["tile"] == shape[0] and d["tiles"] * shape[0] == d["slots"], d
            b.use_graph(graph)
            for i, sc in enumerate(ens):
                b.upload_state(i, sc.pos4)
            for steps, refine in plans:
                b.substep(steps, refine)
            results.append(([b.state(i) for i in range(len(ens))], b.counters()))
        print(f"shape {shape}: {d['groups']} groups, {d['streams']} streams")
    base, base_counters = results[0]
    for (states, counters), shape in zip(results[1:], SHAPES[1:]):
        assert np.array_equal(counters, base_counters), shape
        for i in range(len(ens)):
            for key in STATE:
                assert np.array_equal(_bits(states[i][key]), _bits(base[i][key])), (shape, i, key)
    # and the words of a lone sim
    i = int(np.argmax([sc.n for sc in ens]))
    with fa.Sim(ens[i].params) as s:
        s.upload_state(ens[i].pos4)
        for steps, refine in plans:
            s.substeps(steps[i], 1, refine[i])
        lone = s.state()
    for key in STATE:
        assert np.array_equal(_bits(lone[key]), _bits(base[i][key])), key


def test_the_autotuner_picks_a_valid_shape_and_keeps_results():
    ens = _ensemble()
    plans = [_plan(ens, k) for k in range(2)]
    with fa.Batch([sc.params for sc in ens]) as b:
        for i, sc in enumerate(ens):
            b.upload_state(i, sc.pos4)
        cap = b.autotune(*plans[0], trials=1)
        t = b.tuning()
        print("autotuned shape:", t)
        assert t["tile"] in (64, 128, 256) and t["stage_cap"] == cap and 1 <= t["streams"] <= 4
        assert not b.counters().any()
        for steps, refine in plans:
            b.substep(steps, refine)
        tuned = [b.state(i) for i in range(len(ens))]
    with fa.Batch([sc.params for sc in ens]) as b:
        for i, sc in enumerate(ens):
            b.upload_state(i, sc.pos4)
        for steps, refine in plans:
            b.substep(steps, refine)
        for i in range(len(ens)):
            st = b.state(i)
            for key in STATE:
                assert np.array_equal(_bits(st[key]), _bits(tuned[i][key])), (i, key)
