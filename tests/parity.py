"""Shared helpers of the parity tests: run the HIP step (through the C ABI)
and the oracle on the same seeded input and compare.

Error measure (the tolerance in every FP test): the normwise relative error
of one quantity over all particles, ||gpu - ref||_2 / ||ref||_2, per vector
component set. Quantities clamped at zero (kappa, pressure) have no
meaningful pointwise relative error, so the normwise one is used throughout.
"""
from __future__ import annotations

import numpy as np

import fluid_amd as fa
from oracle import OracleSim

STATE_KEYS = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")
FIELD_KEYS = ("density", "alpha", "kd", "pressure", "rho_near", "accel", "xsph", "wall_grad")
FP32_TOL = 1e-5   # per-substep bar against the FP32 oracle


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0.0:
        return float(num)
    return float(num / den)


def run_gpu(scene, steps, refine, n_sub=1, tracers=None, nbr_cap=0, ids=False, density_only=False):
    """Runs `steps` (a list of Step blocks, one substep each) on the GPU."""
    with fa.Sim(scene.params, tracers=0 if tracers is None else len(tracers), nbr_cap=nbr_cap) as s:
        s.upload_state(scene.pos4)
        if ids:
            s.track_ids(True)
        for st in steps:
            if density_only:
                s.density_only(st)
            else:
                s.substeps(st, n_sub, refine)
        out = s.state()
        out["grid"] = s.grid()
        out["fields"] = {k: s.field(k) for k in FIELD_KEYS}
        out["counters"] = s.counters()
        if ids:
            out["ids"] = s.ids()
        return out


def run_oracle(scene, steps, refine, n_sub=1, precision=32, nbr_cap=96, density_only=False):
    o = OracleSim(scene.params, nbr_cap=nbr_cap, precision=precision)
    o.set_state(scene.pos4)
    for st in steps:
        for _ in range(n_sub):
            if density_only:
                o.density_only(st)
            else:
                o.substep(st, refine)
    out = o.state()
    out["grid"] = o.grid()
    out["fields"] = {k: o.field(k) for k in FIELD_KEYS}
    out["counters"] = o.counters()
    if precision == 64:
        out["f64"] = o.state_f64()
    return out


def compare(gpu, ref, keys=STATE_KEYS, fields=FIELD_KEYS):
    errs = {}
    for k in keys:
        errs[k] = rel_err(gpu[k], ref[k])
    for k in fields:
        errs[k] = rel_err(gpu["fields"][k], ref["fields"][k])
    return errs
