"""Seeded synthetic scenes of the benchmark configurations.

Every scene is a jittered lattice block in a box centred on the origin
(walls at +-half, as the reference's walls sit at -(box_min + cell),
sim_solve.wgsl:237-239). Spacing d; smoothing length h = h_factor * d with
support 2h = grid cell (sim.rs:12-32, render.rs:1477-1478). The configs'
"spacing 0.5h" reads h as the support radius, so their default h_factor
is 1.0 (support 2d); the reference's own default is 1.2 (render.rs:1477).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .sim import Grid, pack_sim_params, pack_step, pack_tracer, refine_passes, substeps_for

FRAME_DT = 1.0 / 120.0   # render.rs:81 NOMINAL_FRAME; the configs' dt
GRAVITY = 9.81


@dataclass
class Scene:
    name: str
    half: tuple
    spacing: float
    h: float
    rho0: float
    lattice: tuple
    pos4: np.ndarray
    grid: Grid
    params: bytes
    tracers: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.uint32))
    phase: float = 0.0
    seed: int = 0
    # A random-walk tilt: the angle of each frame (None: the sinusoid).
    walk: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.pos4.shape[0])

    def gravity(self, t: float):
        """g(t) = 9.81 (sin th, -cos th, 0): th = (pi/4) sin(2 pi t / 2 s +
        phase), or with a random walk th of the frame that holds t."""
        if self.walk is not None:
            k = min(int(t / FRAME_DT), len(self.walk) - 1)
            th = float(self.walk[k])
        else:
            th = (math.pi / 4.0) * math.sin(2.0 * math.pi * t / 2.0 + self.phase)
        return (GRAVITY * math.sin(th), -GRAVITY * math.cos(th), 0.0)

    def schedule(self, v_max: float = 0.0, cap: int = 16, frame_dt: float = FRAME_DT):
        """(substeps, dt_sub, refine, v_clamp) of a frame: render.rs:3090-3094, 3169-3171."""
        n = max(1, substeps_for(frame_dt, v_max, self.spacing, cap))
        dt_sub = frame_dt / n
        return n, dt_sub, refine_passes(dt_sub), 0.4 * self.spacing / dt_sub

    def step_block(self, t: float, dt_sub: float, v_clamp: float, seed: int = 0) -> bytes:
        return pack_step(force=self.gravity(t), dt=dt_sub, v_clamp=v_clamp, seed=seed)


def _lattice(lattice, d, corner, rng, jitter):
    nx, ny, nz = lattice
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([corner[0] + (i + 0.5) * d, corner[1] + (j + 0.5) * d, corner[2] + (k + 0.5) * d], -1)
    pos = pos.reshape(-1, 3).astype(np.float64)
    if jitter > 0:
        pos += rng.uniform(-jitter * d, jitter * d, size=pos.shape)
    out = np.zeros((pos.shape[0], 4), np.float32)
    out[:, :3] = pos.astype(np.float32)
    return out


def _tracers(count, half, fill_height, rng):
    if count == 0:
        return np.zeros((0, 2), np.uint32)
    lo = np.array([-half[0] + 1e-3, -half[1] + 1e-3, -half[2] + 1e-3])
    hi = np.array([half[0] - 1e-3, -half[1] + fill_height, half[2] - 1e-3])
    pts = rng.uniform(lo, hi, size=(count, 3)).astype(np.float32)
    return np.stack([pack_tracer(p, 0.0, (half[0], half[1])) for p in pts]).astype(np.uint32)


def block_scene(name, lattice, spacing, half, place="floor", jitter=0.01, seed=0, h_factor=1.0,
                rho0=1000.0, tracers=0) -> Scene:
    """A lattice block of `lattice` particles at `spacing`; `place` is "floor"
    (centred in x and z, resting on the floor) or "corner" (lower-left of the
    x-y face, full depth from the back wall)."""
    rng = np.random.default_rng(seed)
    d = float(spacing)
    if place == "floor":
        corner = (-0.5 * lattice[0] * d, -half[1], -0.5 * lattice[2] * d)
    elif place == "corner":
        corner = (-half[0], -half[1], -half[2])
    else:
        raise ValueError(place)
    pos4 = _lattice(lattice, d, corner, rng, jitter)
    h = h_factor * d
    grid = Grid.new(half, 2.0 * h)
    mass = rho0 * d * d * d
    params = pack_sim_params(grid, pos4.shape[0], h, mass, rho0)
    phase = float(rng.uniform(0.0, 2.0 * math.pi))
    tr = _tracers(tracers, half, lattice[1] * d, rng)
    return Scene(name, tuple(half), d, h, rho0, tuple(lattice), pos4, grid, params, tr, phase, seed)


def dam_break(seed=0, h_factor=1.0, rho0=1000.0, tracers=0) -> Scene:
    """16,384 particles (32 x 32 x 16) at d = 1/64 in the lower-left quarter
    of a (1, 1, 0.25) box: the plain dam-break configuration."""
    return block_scene("dam_break_16k", (32, 32, 16), 1.0 / 64.0, (0.5, 0.5, 0.125), place="corner",
                       seed=seed, h_factor=h_factor, rho0=rho0, tracers=tracers)


def block_1m(seed=0, h_factor=1.0, tracers=4096) -> Scene:
    """1,048,576 particles (128 x 64 x 128) filling 40% of the unit box, on
    the floor, with 4,096 tracers and a sinusoidal tilt."""
    n = 128 * 64 * 128
    d = (0.4 / n) ** (1.0 / 3.0)
    return block_scene("block_1m", (128, 64, 128), d, (0.5, 0.5, 0.5), seed=seed, h_factor=h_factor,
                       tracers=tracers)


def tilt_walk(frames: int, sigma: float, seed: int) -> np.ndarray:
    """A seeded random-walk tilt, one angle a frame: th_0 = 0, th_k+1 = th_k +
    N(0, sigma^2)."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x7417]))
    return np.concatenate([[0.0], np.cumsum(rng.normal(0.0, sigma, frames - 1))])


def block_8m(seed=0, h_factor=1.0, frames=10000) -> Scene:
    """8,388,608 particles (256 x 128 x 256) filling 30% of the unit box,
    with a random-walk tilt (step sigma 0.05 rad a frame)."""
    n = 256 * 128 * 256
    d = (0.3 / n) ** (1.0 / 3.0)
    sc = block_scene("block_8m", (256, 128, 256), d, (0.5, 0.5, 0.5), seed=seed, h_factor=h_factor)
    sc.walk = tilt_walk(frames, 0.05, seed)
    return sc


def small_block(lattice=(8, 8, 8), seed=0, h_factor=1.2, spacing=None, half=(0.1, 0.1, 0.1), jitter=0.1,
                place="floor", rho0=998.2, tracers=0) -> Scene:
    """A small test scene at the reference's own defaults (h = 1.2 d, rho0 =
    998.2, +-10% jitter: sim.rs:6, 85, render.rs:1477)."""
    d = spacing if spacing is not None else 0.01
    return block_scene(f"small_{lattice[0]}x{lattice[1]}x{lattice[2]}", lattice, d, half, place=place,
                       jitter=jitter, seed=seed, h_factor=h_factor, rho0=rho0, tracers=tracers)


def fill_block(name, n, fill, half=(0.5, 0.5, 0.5), seed=0, h_factor=1.0, rho0=1000.0, jitter=0.01) -> Scene:
    """Exactly n particles filling `fill` of the box volume: a 2:1:2 lattice
    block on the floor (its top layer partly filled), spacing
    d = (fill * volume / n)^(1/3)."""
    vol = 8.0 * half[0] * half[1] * half[2]
    d = (fill * vol / n) ** (1.0 / 3.0)
    nx = max(1, int(round((2.0 * n) ** (1.0 / 3.0))))
    ny = int(math.ceil(n / (nx * nx)))
    rng = np.random.default_rng(seed)
    corner = (-0.5 * nx * d, -half[1], -0.5 * nx * d)
    pos4 = _lattice((nx, ny, nx), d, corner, rng, jitter)
    keep = np.sort(np.argsort(pos4[:, 1], kind="stable")[:n])
    pos4 = np.ascontiguousarray(pos4[keep])
    h = h_factor * d
    grid = Grid.new(half, 2.0 * h)
    params = pack_sim_params(grid, n, h, rho0 * d * d * d, rho0)
    phase = float(rng.uniform(0.0, 2.0 * math.pi))
    return Scene(name, tuple(half), d, h, rho0, (nx, ny, nx), pos4, grid, params, np.zeros((0, 2), np.uint32),
                 phase, seed)


def ensemble(scenes=256, seed=0, n_min=2048, n_max=65536, fill=(0.2, 0.5), h_factor=1.0) -> list:
    """The batched configuration: `scenes` independent unit boxes, particle
    counts log-uniform in [n_min, n_max], fill fractions uniform in `fill`,
    each with its own seeded tilt phase."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=scenes))).astype(int)
    fills = rng.uniform(fill[0], fill[1], size=scenes)
    return [fill_block(f"ensemble_{i}", int(counts[i]), float(fills[i]), seed=seed * 100003 + i, h_factor=h_factor)
            for i in range(scenes)]
