"""Creation refuses what the step cannot hold, with the reason, before it
allocates: a neighbour cap that is not a multiple of 4 in [4, 256], a list
whose 32-bit offsets would overflow (cap x slots >= 2^32), a grid under 2
cells on an axis, a non-positive smoothing length. A sim and a batch both.
"""
import struct

import pytest

import fluid_amd as fa

pytestmark = pytest.mark.gpu


def _params(sc, count=None, dims=None, h=None):
    """sc's SimParams block (sim.rs:325-350) with the grid dims (bytes 16-27),
    the count (28-31) or h (32-35) replaced."""
    p = bytearray(sc.params)
    if count is not None:
        p[28:32] = struct.pack("<I", count)
    if dims is not None:
        p[16:28] = struct.pack("<3I", *dims)
    if h is not None:
        p[32:36] = struct.pack("<f", h)
    return bytes(p)


@pytest.mark.parametrize("cap", [6, 260])
def test_a_cap_off_the_quad_grid_is_refused(cap):
    sc = fa.scenes.small_block((4, 4, 4), seed=51)
    with pytest.raises(RuntimeError, match="multiple of 4"):
        fa.Sim(sc.params, nbr_cap=cap)
    with pytest.raises(RuntimeError, match="multiple of 4"):
        fa.Batch([sc.params], nbr_cap=cap)


@pytest.mark.parametrize("count,cap", [(45_000_000, 96), (16_777_216, 256)])
def test_a_list_past_32_bit_offsets_is_refused(count, cap):
    sc = fa.scenes.small_block((4, 4, 4), seed=52)
    big = _params(sc, count=count)
    with pytest.raises(RuntimeError, match="exceeds 2\\^32"):
        fa.Sim(big, nbr_cap=cap)
    with pytest.raises(RuntimeError, match="exceeds 2\\^32"):
        fa.Batch([big], nbr_cap=cap)


def test_a_flat_grid_is_refused():
    sc = fa.scenes.small_block((4, 4, 4), seed=53)
    flat = _params(sc, dims=(8, 1, 8))
    with pytest.raises(RuntimeError, match="at least 2"):
        fa.Sim(flat)
    with pytest.raises(RuntimeError, match="at least 2"):
        fa.Batch([sc.params, flat])


@pytest.mark.parametrize("h", [0.0, -0.01, float("nan")])
def test_a_non_positive_smoothing_length_is_refused(h):
    from fluid_amd import adjoint as fad
    sc = fa.scenes.small_block((4, 4, 4), seed=54)
    bad = _params(sc, h=h)
    with pytest.raises(RuntimeError, match="must be positive"):
        fa.Sim(bad)
    with pytest.raises(RuntimeError, match="must be positive"):
        fa.Batch([sc.params, bad])
    with pytest.raises(RuntimeError, match="must be positive"):
        fad.Adjoint(bad, 2, 1)


def test_the_largest_refusal_leaves_the_device_usable():
    sc = fa.scenes.small_block((4, 4, 4), seed=55)
    with pytest.raises(RuntimeError):
        fa.Sim(_params(sc, count=45_000_000))
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        n, dt, r, vc = sc.schedule()
        s.substeps(sc.step_block(0.0, dt, vc), 1, r)
        assert s.state()["pos"].shape == (sc.n, 4)
