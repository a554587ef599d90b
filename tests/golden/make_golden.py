"""Writes the step fixtures of tests/golden/: seeded tiny scenes, their
substep inputs, and what the oracle makes of them.

The reference holds no vectors for its step (it tests bounds only, and it
cannot run here: Rust + WGSL on Metal), so these fixtures are the oracle's
own answers, frozen. They catch any later change to the oracle or the HIP
step that moves a bit; the oracle's correctness is pinned separately against
closed forms, quadrature, brute force and invariants.

Each fixture `step_<case>.npz` (no pickled objects) holds:
- params (u8[64]), steps (u8[k, 144]), refine (i32): the SimParams block,
  one Step block per substep, the density passes of each substep;
- pos4_in (f32[n, 4]): the resting set before the first substep;
- from the FP32 oracle after the last substep: pos, vel, prev_vel,
  prev_pressure, temperature (f32), perm, nbr_n (u32), counters (u64[2]:
  clamps, neighbour overflows), density and pressure (f32, last substep);
- from the FP64 oracle after the last substep: pos64 and vel64 (f64).

Run from the repository root: python3 tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "fluid-simulation-app_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from fluid_amd import scenes  # noqa: E402
from fluid_amd.sim import pack_step  # noqa: E402
from oracle import OracleSim  # noqa: E402


def _steps(sc, k, omega=(0.0, 0.0, 0.0), domega=(0.0, 0.0, 0.0), touch=None):
    """k substeps of the scene's own schedule, gravity tilting with t; the
    first substep carries the rotation and the finger, if any."""
    _, dt, refine, v_clamp = sc.schedule()
    out = []
    for i in range(k):
        kw = {}
        if i == 0:
            kw = dict(omega=omega, domega=domega)
            if touch is not None:
                kw.update(touches=[touch[:4]], touch_n=1, touch_r=touch[4])
        out.append(pack_step(force=sc.gravity(i * dt), dt=dt, v_clamp=v_clamp, seed=i, **kw))
    return out, refine


def cases():
    """(name, scene, steps, refine) of every fixture."""
    lone = scenes.small_block((1, 1, 1), seed=11)
    s64 = scenes.small_block((4, 4, 4), seed=12)
    s512 = scenes.small_block((8, 8, 8), seed=13)
    dam = scenes.block_scene("dam_2k", (16, 16, 8), 1.0 / 48.0, (0.25, 0.25, 0.0833), place="corner",
                             seed=14, jitter=0.02)
    out = []
    st, r = _steps(lone, 1)
    out.append(("lone_1", lone, st, r))
    st, r = _steps(s64, 1)
    out.append(("small_64", s64, st, r))
    st, r = _steps(s512, 2, omega=(0.0, 0.0, 1.5), domega=(0.0, 0.0, 4.0))
    out.append(("small_512_rotating", s512, st, r))
    st, r = _steps(dam, 3, touch=(-0.2, -0.2, 0.5, 0.0, 0.05))
    out.append(("dam_2k_touched", dam, st, r))
    return out


def run(sc, steps, refine, precision):
    o = OracleSim(sc.params, precision=precision)
    o.set_state(sc.pos4)
    for st in steps:
        o.substep(st, refine)
    return o


def fixture(sc, steps, refine) -> dict:
    o = run(sc, steps, refine, 32)
    s = o.state()
    g = o.grid()
    out = dict(params=np.frombuffer(sc.params, np.uint8).copy(),
               steps=np.stack([np.frombuffer(b, np.uint8) for b in steps]),
               refine=np.int32(refine), pos4_in=np.ascontiguousarray(sc.pos4, np.float32),
               perm=g["perm"], nbr_n=g["nbr_n"], counters=np.array(o.counters(), np.uint64),
               density=o.field("density").astype(np.float32), pressure=o.field("pressure").astype(np.float32))
    for k in ("pos", "vel", "prev_vel", "prev_pressure", "temperature"):
        out[k] = s[k]
    o.close()
    o64 = run(sc, steps, refine, 64)
    s64 = o64.state_f64()
    out["pos64"], out["vel64"] = s64["pos"], s64["vel"]
    o64.close()
    return out


def main():
    for name, sc, steps, refine in cases():
        path = os.path.join(HERE, f"step_{name}.npz")
        np.savez_compressed(path, **fixture(sc, steps, refine))
        print(f"{path}: {sc.n} particles, {len(steps)} substeps, refine {refine}, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
