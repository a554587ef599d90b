"""The HIP step against the committed step fixtures (tests/golden/),
through the C ABI.

Every state word, the permutation, the neighbour counts, the counters and
the last substep's density and pressure equal the FP32 fixture bit for bit.
Against the FP64 fixture, positions and velocities stay within the
per-substep bar (normwise relative error 1e-5) times the substeps run. The
batched step, all fixtures as the scenes of one batch, gives the same words.
"""
import numpy as np
import pytest

import fluid_amd as fa
from golden_io import STATE, bits, load, names
from parity import FP32_TOL, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", names())
def test_hip_step_reproduces_the_fixture(name):
    f = load(name)
    print(f"golden {name}: {f['pos4_in'].shape[0]} particles, {len(f['step_bytes'])} substeps", flush=True)
    with fa.Sim(f["params_bytes"]) as s:
        s.upload_state(f["pos4_in"])
        for st in f["step_bytes"]:
            s.substeps(st, 1, f["refine"])
        out = s.state()
        g = s.grid()
        counters = s.counters()
        fields = {k: s.field(k) for k in ("density", "pressure")}
    for k in STATE:
        assert np.array_equal(bits(out[k]), bits(f[k])), k
    assert np.array_equal(g["perm"], f["perm"])
    assert np.array_equal(g["nbr_n"], f["nbr_n"])
    assert tuple(int(c) for c in counters) == tuple(int(c) for c in f["counters"])
    for k, v in fields.items():
        assert np.array_equal(bits(v), bits(f[k])), k
    bar = FP32_TOL * len(f["step_bytes"])
    for k in ("pos", "vel"):
        e = rel_err(out[k][:, :3], f[k + "64"][:, :3])
        print(f"  {k} vs fp64 fixture: {e:.3e} (bar {bar:.0e})")
        assert e <= bar, k


def test_one_batch_reproduces_every_fixture():
    """All fixtures as the scenes of one batch: each substep, a scene past
    its own substeps gets a Step with dt = 0 and sits out. Every scene's
    state words and counters equal its fixture's."""
    fs = [load(name) for name in names()]
    rounds = max(len(f["step_bytes"]) for f in fs)
    idle = fa.pack_step(force=(0.0, -9.81, 0.0), dt=0.0, v_clamp=1.0)
    print(f"golden batch: {len(fs)} scenes, {rounds} substeps", flush=True)
    with fa.Batch([f["params_bytes"] for f in fs]) as b:
        for i, f in enumerate(fs):
            b.upload_state(i, f["pos4_in"])
        for k in range(rounds):
            steps = [f["step_bytes"][k] if k < len(f["step_bytes"]) else idle for f in fs]
            b.substep(steps, [f["refine"] for f in fs])
        counters = b.counters()
        for i, f in enumerate(fs):
            st = b.state(i)
            for key in STATE:
                assert np.array_equal(bits(st[key]), bits(f[key])), (names()[i], key)
            assert tuple(int(c) for c in counters[i]) == tuple(int(c) for c in f["counters"]), names()[i]
