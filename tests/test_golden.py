"""The oracle against the committed step fixtures (tests/golden/).

The fixtures freeze the oracle's answers on four seeded tiny scenes (a lone
particle; 64 particles; 512 particles under rotation; 2,048 particles in a
corner with a finger, three substeps). The FP32 oracle must reproduce every
word; the FP64 oracle every value to 1e-13 relative. The scene builders and
the Step packer must reproduce the fixtures' inputs.
"""
import numpy as np
import pytest

from golden.make_golden import cases
from golden_io import STATE, bits, load, names
from oracle import OracleSim


def test_every_case_has_a_fixture():
    assert names() == sorted(c[0] for c in cases())


@pytest.mark.parametrize("name", names())
def test_fp32_oracle_reproduces_the_fixture_bit_for_bit(name):
    f = load(name)
    o = OracleSim(f["params_bytes"], precision=32)
    o.set_state(f["pos4_in"])
    for st in f["step_bytes"]:
        o.substep(st, f["refine"])
    s, g = o.state(), o.grid()
    for k in STATE:
        assert np.array_equal(bits(s[k]), bits(f[k])), k
    assert np.array_equal(g["perm"], f["perm"])
    assert np.array_equal(g["nbr_n"], f["nbr_n"])
    assert tuple(o.counters()) == tuple(int(c) for c in f["counters"])
    for k in ("density", "pressure"):
        assert np.array_equal(bits(o.field(k).astype(np.float32)), bits(f[k])), k
    o.close()


@pytest.mark.parametrize("name", names())
def test_fp64_oracle_reproduces_the_fixture(name):
    f = load(name)
    o = OracleSim(f["params_bytes"], precision=64)
    o.set_state(f["pos4_in"])
    for st in f["step_bytes"]:
        o.substep(st, f["refine"])
    s = o.state_f64()
    o.close()
    for k in ("pos", "vel"):
        np.testing.assert_allclose(s[k], f[k + "64"], rtol=1e-13, atol=1e-15, err_msg=k)


def test_fixture_inputs_are_what_the_builders_make():
    for name, sc, steps, refine in cases():
        f = load(name)
        assert f["params_bytes"] == sc.params, name
        assert f["step_bytes"] == steps, name
        assert f["refine"] == refine, name
        assert np.array_equal(bits(f["pos4_in"]), bits(sc.pos4)), name


def test_fixtures_hold_the_physics_they_were_chosen_for():
    """A lone particle has no neighbours and reads m W(0) plus the wall
    fill (positive, below rho0); the finger case clamps a speed; the 512
    case holds more than 48 entries somewhere (the LDS walk list's length)."""
    lone = load("lone_1")
    assert lone["nbr_n"].tolist() == [0]
    assert 0.0 < float(lone["density"][0]) < 998.2
    assert int(load("dam_2k_touched")["counters"][0]) >= 1
    assert int(load("small_512_rotating")["nbr_n"].max()) > 48
