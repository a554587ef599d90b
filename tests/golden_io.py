"""Loading of the step fixtures under tests/golden/ (written by
tests/golden/make_golden.py)."""
from __future__ import annotations

import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATE = ("pos", "vel", "prev_vel", "prev_pressure", "temperature")


def names():
    return sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "step_*.npz")))


def load(name: str) -> dict:
    with np.load(os.path.join(GOLDEN, f"step_{name}.npz")) as d:
        out = {k: d[k] for k in d.files}
    out["params_bytes"] = out["params"].tobytes()
    out["step_bytes"] = [row.tobytes() for row in out["steps"]]
    out["refine"] = int(out["refine"])
    return out


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)
