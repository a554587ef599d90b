"""Kernel timing through the C ABI (fluid_sim_profile_sampled): every
launch of a class, or one in every `stride`, and the step's bits either
way."""
import numpy as np
import pytest

import fluid_amd as fa

pytestmark = pytest.mark.gpu


def _run(sc, steps, refine, stride):
    s = fa.Sim(sc.params)
    try:
        s.upload_state(sc.pos4)
        s.profile(["den_kappa", "density_div"], stride=stride)
        for st in steps:
            s.substeps(st, 1, refine)
        return s.profile_read(), s.state()
    finally:
        s.close()


def test_sampled_timing_times_one_launch_in_stride_and_keeps_the_bits():
    sc = fa.scenes.dam_break(seed=0)
    n_sub, dt_sub, refine, v_clamp = sc.schedule()
    steps = [sc.step_block(k * dt_sub, dt_sub, v_clamp) for k in range(6)]
    every, every_state = _run(sc, steps, refine, 1)
    some, some_state = _run(sc, steps, refine, 3)
    launches = len(steps) * refine
    assert every["den_kappa"][1] == launches
    assert some["den_kappa"][1] == -(-launches // 3)
    # one density_div a substep: launches 0 and 3 of 6
    assert every["density_div"][1] == len(steps)
    assert some["density_div"][1] == 2
    assert some["den_kappa"][0] > 0.0 and some["density_div"][0] > 0.0
    for k in every_state:
        assert np.array_equal(every_state[k].view(np.uint32), some_state[k].view(np.uint32)), k
