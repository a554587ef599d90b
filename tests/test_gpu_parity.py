"""Parity of the HIP step with the oracle, through the C ABI.

Integer work (cell ids, histogram, scan, permutation, neighbour lists,
counters, packed tracers) is compared bit for bit. Floating-point fields are
held to the normwise relative error of tests/parity.py: <= 1e-5 per substep
against the FP32 oracle; the error against the FP64 oracle is printed.
"""
import numpy as np
import pytest

import fluid_amd as fa
from oracle import OracleSim
from parity import FP32_TOL, compare, rel_err, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _scene_small(seed=3):
    return fa.scenes.small_block((12, 12, 12), seed=seed)


def _step(scene, t=0.0):
    n, dt, refine, v_clamp = scene.schedule()
    return scene.step_block(t, dt, v_clamp), refine


@pytest.mark.parametrize("cells", [1, 7, 2047, 2048, 2049, 100_003, 4_000_037])
def test_scan_matches_exclusive_cumsum(cells):
    rng = np.random.default_rng(cells)
    counts = rng.integers(0, 40, size=cells, dtype=np.uint32)
    counts[rng.integers(0, cells, size=max(1, cells // 10))] = 0
    starts = fa.scan_u32(counts)
    expect = np.zeros(cells + 1, np.uint64)
    expect[1:] = np.cumsum(counts, dtype=np.uint64)
    assert np.array_equal(starts.astype(np.uint64), expect)


@pytest.mark.parametrize("maker", ["small", "dam"])
def test_grid_build_is_bit_exact(maker):
    sc = _scene_small() if maker == "small" else fa.scenes.dam_break(seed=1)
    step, _ = _step(sc)
    g = run_gpu(sc, [step], 0, density_only=True)
    o = run_oracle(sc, [step], 0, density_only=True)
    for k in ("cellid", "counts", "starts", "perm"):
        assert np.array_equal(g["grid"][k], o["grid"][k]), k
    assert g["grid"]["starts"][-1] == sc.n
    # the working set is the resting set in permutation order, bit for bit
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        s.density_only(step)
        wp = s.field("work_pos")
        perm = s.grid()["perm"]
    assert np.array_equal(wp.view(np.uint32), sc.pos4[perm].view(np.uint32))


def test_density_sweep_matches_oracle():
    sc = fa.scenes.dam_break(seed=2, h_factor=1.2, rho0=998.2)
    step, _ = _step(sc)
    g = run_gpu(sc, [step], 0, density_only=True)
    o = run_oracle(sc, [step], 0, density_only=True)
    o64 = run_oracle(sc, [step], 0, precision=64, density_only=True)
    assert np.array_equal(g["grid"]["nbr_n"], o["grid"]["nbr_n"])
    for k in ("density", "alpha", "kd", "rho_near", "wall_grad"):
        e32 = rel_err(g["fields"][k], o["fields"][k])
        e64 = rel_err(g["fields"][k], o64["fields"][k])
        print(f"{k}: vs fp32 oracle {e32:.3e}, vs fp64 oracle {e64:.3e}")
        assert e32 <= FP32_TOL, k


@pytest.mark.parametrize("maker", ["small", "dam"])
def test_one_substep_matches_oracle(maker):
    sc = _scene_small() if maker == "small" else fa.scenes.dam_break(seed=4)
    step, refine = _step(sc)
    g = run_gpu(sc, [step], refine)
    o = run_oracle(sc, [step], refine)
    o64 = run_oracle(sc, [step], refine, precision=64)
    assert np.array_equal(g["grid"]["perm"], o["grid"]["perm"])
    errs = compare(g, o)
    errs64 = compare(g, o64)
    for k in errs:
        print(f"{k}: vs fp32 {errs[k]:.3e}  vs fp64 {errs64[k]:.3e}")
    bad = {k: v for k, v in errs.items() if v > FP32_TOL}
    assert not bad, bad
    assert g["counters"] == o["counters"]


def test_ten_substeps_keep_every_particle_and_word():
    sc = fa.scenes.dam_break(seed=5)
    step, refine = _step(sc)
    steps = [step] * 10
    g = run_gpu(sc, steps, refine, ids=True)
    o = run_oracle(sc, steps, refine)
    # the same particle in the same slot, then the words themselves
    assert np.array_equal(g["ids"], o["ids"])
    for k in ("pos", "vel"):
        err = rel_err(g[k], o[k])
        print(k, "drift after 10 substeps vs fp32 oracle", err)
        assert err == 0.0, k


def test_200_substeps_stay_bit_exact_with_the_fp32_oracle():
    """The dam break (16,384 particles) over 200 substeps of its schedule,
    gravity tilting with t: every state word equals the FP32 oracle's, so
    the GPU's drift from the FP64 oracle is the FP32 method's own (printed,
    by particle id: the flow is chaotic, and by 50 substeps single
    trajectories have parted by a few percent of the box)."""
    sc = fa.scenes.dam_break(seed=5)
    n, dt, refine, vc = sc.schedule()
    steps = [sc.step_block(k * dt, dt, vc) for k in range(200)]
    print(f"dam break {sc.n} particles, 200 substeps, R = {refine}", flush=True)
    g = run_gpu(sc, steps, refine, ids=True)
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    for st in steps:
        o.substep(st, refine)
    ost = o.state()
    assert np.array_equal(g["ids"], ost["ids"])
    for k in ("pos", "vel", "prev_vel", "prev_pressure", "temperature"):
        assert np.array_equal(_bits(g[k]), _bits(ost[k])), k
    o64 = OracleSim(sc.params, precision=64)
    o64.set_state(sc.pos4)
    for st in steps:
        o64.substep(st, refine)
    f64, ids64 = o64.state_f64(), o64.state()["ids"]
    ia, ib = np.argsort(g["ids"]), np.argsort(ids64)
    for k in ("pos", "vel"):
        print(k, "after 200 substeps vs fp64 oracle, by particle:", rel_err(g[k][ia, :3], f64[k][ib, :3]))


def test_tracers_are_bit_exact():
    sc = fa.scenes.small_block((10, 10, 10), seed=9, tracers=2000)
    step, refine = _step(sc)
    frame = fa.pack_step(force=(0.0, -9.81, 0.0), dt=1.0 / 120.0, seed=5)
    with fa.Sim(sc.params, tracers=len(sc.tracers)) as s:
        s.upload_state(sc.pos4)
        s.upload_tracers(sc.tracers)
        s.substeps(step, 2, refine)
        s.tracers(frame)
        gt = s.download_tracers()
        st = s.state()
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    o.substep(step, refine)
    o.substep(step, refine)
    ot = o.tracers(frame, sc.tracers)
    assert np.array_equal(st["pos"].view(np.uint32), o.state()["pos"].view(np.uint32))
    assert np.array_equal(gt.reshape(-1), ot.reshape(-1))


def test_stats_match_oracle():
    sc = _scene_small(seed=11)
    step, refine = _step(sc)
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        s.substeps(step, 2, refine)
        s.reduce_stats()
        gs = s.read_stats()
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    o.substep(step, refine)
    o.substep(step, refine)
    os_ = o.reduce_stats()
    assert abs(gs[0] - os_[0]) <= 1e-5 * max(abs(os_[0]), 1e-6)
    np.testing.assert_allclose(gs[1:], os_[1:], rtol=1e-5, atol=1e-6)


def test_run_twice_is_bitwise_deterministic():
    sc = fa.scenes.dam_break(seed=6)
    step, refine = _step(sc)
    a = run_gpu(sc, [step] * 3, refine)
    b = run_gpu(sc, [step] * 3, refine)
    for k in ("pos", "vel", "prev_vel", "prev_pressure", "temperature"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def test_single_particle_and_dt_zero():
    sc = fa.scenes.small_block((1, 1, 1), seed=1)
    step, refine = _step(sc)
    g = run_gpu(sc, [step], refine)
    o = run_oracle(sc, [step], refine)
    for k, v in compare(g, o).items():
        assert v <= FP32_TOL, k
    zero = fa.pack_step(force=(0, -9.81, 0), dt=0.0, v_clamp=1.0)
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        s.substeps(zero, 3, 5)
        assert np.array_equal(s.state()["pos"], sc.pos4)


@pytest.mark.parametrize("lds", [None, 8192])
def test_neighbour_cap_overflow_is_counted_like_the_oracle(lds, monkeypatch):
    """A cap of 8, below the list's fixed rows, in staged tiles and (8 KB of
    LDS) unstaged ones, whose walk writes the list directly."""
    if lds:
        monkeypatch.setenv("FLUID_LDS_BYTES", str(lds))
    sc = _scene_small(seed=12)
    step, refine = _step(sc)
    g = run_gpu(sc, [step], refine, nbr_cap=8)
    o = run_oracle(sc, [step], refine, nbr_cap=8)
    assert g["counters"] == o["counters"] and g["counters"][1] > 0
    for k, v in compare(g, o).items():
        assert v <= FP32_TOL, k


def test_touches_and_rotation_match_oracle():
    sc = _scene_small(seed=13)
    n, dt, refine, v_clamp = sc.schedule()
    step = fa.pack_step(force=(1.0, -9.81, 0.5), omega=(0.3, -0.2, 0.9), domega=(1.0, 0.5, -2.0), dt=dt,
                        v_clamp=v_clamp, touches=[(0.0, -0.05, 0.5, 0.2), (0.02, -0.07, -0.3, 0.1)], touch_n=2,
                        touch_r=0.03)
    g = run_gpu(sc, [step], refine)
    o = run_oracle(sc, [step], refine)
    for k, v in compare(g, o).items():
        assert v <= FP32_TOL, k


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("maker", ["small", "dam"])
def test_one_substep_is_bit_exact_with_the_fp32_oracle(maker):
    """Same operand order, no contraction, correctly rounded sqrt and divide
    on both sides: every state word and field word is identical, and so is
    every neighbour list."""
    sc = _scene_small(seed=21) if maker == "small" else fa.scenes.dam_break(seed=22)
    step, refine = _step(sc)
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        s.substeps(step, 1, refine)
        st = s.state()
        fields = {k: s.field(k) for k in ("density", "alpha", "kd", "pressure", "rho_near", "accel", "xsph")}
        nbr, nbr_n = s.neighbours(), s.grid()["nbr_n"]
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    o.substep(step, refine)
    ost = o.state()
    for k in ("pos", "vel", "prev_vel", "prev_pressure", "temperature"):
        assert np.array_equal(_bits(st[k]), _bits(ost[k])), k
    for k, v in fields.items():
        assert np.array_equal(_bits(v), _bits(o.field(k))), k
    onbr, on = o.nbr(), o.grid()["nbr_n"]
    assert np.array_equal(nbr_n, on)
    for i in range(sc.n):
        assert np.array_equal(nbr[i, : nbr_n[i]], onbr[i, : on[i]]), i


@pytest.mark.parametrize("lds", [8192, 57344])
def test_global_gather_path_matches_staged_path(lds, monkeypatch):
    """Tiles whose neighbour rows exceed the staging capacity gather from
    global memory and their lists hold slots; the results are the same bits.
    FLUID_LDS_BYTES shrinks the capacity (8 KB: no tile staged; 56 KB: some)."""
    sc = fa.scenes.dam_break(seed=23)
    step, refine = _step(sc)
    ref = run_gpu(sc, [step] * 2, refine)
    monkeypatch.setenv("FLUID_LDS_BYTES", str(lds))
    with fa.Sim(sc.params) as s:
        d = s.describe()
        assert d["stage_cap"] < 1600, d
    out = run_gpu(sc, [step] * 2, refine)
    for k in STATE:
        assert np.array_equal(_bits(out[k]), _bits(ref[k])), k
    for k in ("density", "kd", "accel"):
        assert np.array_equal(_bits(out["fields"][k]), _bits(ref["fields"][k])), k
    assert np.array_equal(out["grid"]["nbr_n"], ref["grid"]["nbr_n"])


STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


@pytest.mark.parametrize("h_factor,cap,lds", [(1.5, 96, None), (1.5, 28, None), (1.35, 64, 8192),
                                              (1.5, 96, 8192)])
def test_dense_lists_are_bit_exact_with_the_oracle(h_factor, cap, lds, monkeypatch):
    """Dense scenes (~60 to ~110 neighbours inside the block): the list
    entries of the fixed rows, those packed past them, those past the 48 the
    walk keeps in LDS, and the overflow past the cap, in staged tiles and
    (8 KB of LDS) unstaged ones."""
    if lds:
        monkeypatch.setenv("FLUID_LDS_BYTES", str(lds))
    sc = fa.scenes.small_block((12, 12, 12), seed=31, h_factor=h_factor)
    step, refine = _step(sc)
    with fa.Sim(sc.params, nbr_cap=cap) as s:
        s.upload_state(sc.pos4)
        s.substeps(step, 1, refine)
        st = s.state()
        nbr, nbr_n = s.neighbours(), s.grid()["nbr_n"]
        counters = s.counters()
    o = OracleSim(sc.params, nbr_cap=cap)
    o.set_state(sc.pos4)
    o.substep(step, refine)
    onbr, on = o.nbr(), o.grid()["nbr_n"]
    assert np.array_equal(nbr_n, on)
    assert nbr_n.max() > min(48, cap - 1)  # past the LDS list, or up to a smaller cap
    for i in range(sc.n):
        assert np.array_equal(nbr[i, : nbr_n[i]], onbr[i, : on[i]]), i
    ost = o.state()
    for k in STATE:
        assert np.array_equal(_bits(st[k]), _bits(ost[k])), k
    assert counters == o.counters()


def test_grid_count_taken_by_integrate_follows_uploads_and_density_sweeps():
    """integrate takes the next grid build's count from the positions it
    writes; an upload between substeps, or a density-only sweep, makes the
    next build count again. Each path is held bit for bit to the oracle,
    including the grid of the last build."""
    sc = _scene_small(seed=21)
    step, refine = _step(sc)
    moved = sc.pos4.copy()
    moved[:, 1] += np.float32(0.003)  # every particle shifted, many across a cell face
    with fa.Sim(sc.params) as s:
        s.upload_state(sc.pos4)
        s.substeps(step, 2, refine)        # the second build takes the first integrate's count
        g2 = s.state()
        s.upload_state(moved)              # the count is stale: the next build counts again
        s.substeps(step, 1, refine)
        g3 = s.state()
        grid3 = s.grid()
        s.density_only(step)               # takes integrate's count, leaves none
        s.substeps(step, 1, refine)
        g4 = s.state()
        grid4 = s.grid()
    o = OracleSim(sc.params)
    o.set_state(sc.pos4)
    o.substep(step, refine)
    o.substep(step, refine)
    o2 = o.state()
    o.set_state(moved)
    o.substep(step, refine)
    o3, og3 = o.state(), o.grid()
    o.substep(step, refine)  # density_only then a substep: the substep rebuilds from the same positions
    o4, og4 = o.state(), o.grid()
    for got, want, what in ((g2, o2, "two substeps"), (g3, o3, "after the upload"), (g4, o4, "after density_only")):
        for k in STATE:
            assert np.array_equal(_bits(got[k]), _bits(want[k])), (what, k)
    for got, want in ((grid3, og3), (grid4, og4)):
        for k in ("cellid", "counts", "starts", "perm"):
            assert np.array_equal(got[k], want[k]), k
