"""The adjoint of the DFSPH step behind a torch.autograd.Function.

`Adjoint` wraps the C ABI (include/fluid_sim.h, fluid_adjoint_*): a forward
run of a fixed number of substeps with checkpoints, and the hand-written
reverse sweeps on the GPU. `simulate(adj, v0, mu, gravity)` is the
differentiable function final positions = step^K(x0, v0; mu, g_1..g_K) for
the test harness; torch holds only host tensors here (the state lives in the
library's device buffers).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .sim import _f32p, _u32p, _u8p, pack_step

# Physics settings past the viscosity (fluid_adjoint_set_physics's last four
# values). FULL is the reference's step; VISCOUS switches off near pressure,
# cohesion, adhesion and the XSPH blend.
FULL = dict(k_near=3000.0, tension_scale=1.0, adhesion_scale=1.0, xsph_rate=6.0)
VISCOUS = dict(k_near=0.0, tension_scale=0.0, adhesion_scale=0.0, xsph_rate=0.0)


class Adjoint:
    """One scene, `steps` substeps, a checkpoint every `every` substeps."""

    def __init__(self, params64: bytes, steps: int, every: int = 8, nbr_cap: int = 0, physics=None):
        """`physics`: a dict of k_near, tension_scale, adhesion_scale and
        xsph_rate (FULL, VISCOUS or any mix); VISCOUS when None."""
        self.physics = dict(VISCOUS if physics is None else physics)
        self.L = lib()
        self.h = self.L.fluid_adjoint_create(_u8p(params64), int(steps), int(every), int(nbr_cap))
        if not self.h:
            raise RuntimeError("fluid_adjoint_create failed: " + self.L.fluid_sim_last_error().decode())
        self.n = int(np.frombuffer(params64, np.uint32, 1, 28)[0])
        self.steps = int(steps)
        self.mu = None
        # Counts forward runs: the device tapes belong to the last one, so a
        # backward can tell whether they are still those of its forward.
        self.forwards = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.fluid_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mu(self, mu: float):
        """Viscosity mu with this Adjoint's other physics settings."""
        p = self.physics
        arr = (C.c_float * 5)(float(mu), p["k_near"], p["tension_scale"], p["adhesion_scale"], p["xsph_rate"])
        check(self.L.fluid_adjoint_set_physics(self.h, arr), "adjoint set_physics")
        self.mu = float(mu)

    def forward(self, pos4, vel4, steps144: list, refine):
        blob = b"".join(steps144)
        assert len(blob) == 144 * self.steps
        r = np.ascontiguousarray(refine, np.uint32)
        p = np.ascontiguousarray(pos4, np.float32)
        v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        pos_out = np.empty((self.n, 4), np.float32)
        vel_out = np.empty((self.n, 4), np.float32)
        check(self.L.fluid_adjoint_forward(self.h, _f32p(p), _f32p(v), _u8p(blob), _u32p(r), _f32p(pos_out),
                                           _f32p(vel_out)), "adjoint forward")
        self.forwards += 1
        return pos_out, vel_out

    def backward(self, dpos4, dvel4=None):
        dp = np.ascontiguousarray(dpos4, np.float32)
        dv = None if dvel4 is None else np.ascontiguousarray(dvel4, np.float32)
        g_v0 = np.empty((self.n, 4), np.float32)
        g_x0 = np.empty((self.n, 4), np.float32)
        g_g = np.empty((self.steps, 3), np.float32)
        g_mu = np.empty(1, np.float32)
        check(self.L.fluid_adjoint_backward(self.h, _f32p(dp), _f32p(dv), _f32p(g_v0), _f32p(g_x0), _f32p(g_g),
                                            _f32p(g_mu)), "adjoint backward")
        return g_v0, g_x0, g_g, float(g_mu[0])

    def timing(self):
        a, b = C.c_double(), C.c_double()
        check(self.L.fluid_adjoint_timing(self.h, C.byref(a), C.byref(b)), "adjoint timing")
        return float(a.value), float(b.value)


def steps_for(gravity, dt: float, v_clamp: float, omega=(0.0, 0.0, 0.0), domega=(0.0, 0.0, 0.0)):
    """One Step block per substep from a (K, 3) gravity array, in a frame
    turning at omega (rad/s) with angular acceleration domega."""
    return [pack_step(force=tuple(float(x) for x in g), omega=omega, domega=domega, dt=dt, v_clamp=v_clamp, seed=k)
            for k, g in enumerate(np.asarray(gravity, np.float64))]


def simulate_function():
    """The torch.autograd.Function (built on first use, so importing this
    module does not import torch)."""
    import torch

    class Simulate(torch.autograd.Function):
        """final positions (N, 3) = K substeps of the step from (x0, v0) with
        viscosity mu and per-substep gravity g (K, 3), in a frame turning at
        frame = (omega, domega) (not differentiated)."""

        @staticmethod
        def forward(ctx, v0, mu, gravity, adj, x0, dt, v_clamp, refine, frame=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0))):
            adj.set_mu(float(mu))
            vel4 = np.zeros((adj.n, 4), np.float32)
            vel4[:, :3] = v0.detach().cpu().numpy()
            steps = steps_for(gravity.detach().cpu().numpy(), dt, v_clamp, frame[0], frame[1])
            pos_out, _ = adj.forward(x0, vel4, steps, [refine] * adj.steps)
            ctx.adj = adj
            ctx.forward_id = adj.forwards
            return torch.from_numpy(pos_out[:, :3].copy()).to(v0.dtype)

        @staticmethod
        def backward(ctx, grad_pos):
            adj = ctx.adj
            if adj.forwards != ctx.forward_id:
                # Another forward on this Adjoint replaced the tapes this
                # output was recorded with; its gradient would be that run's.
                raise RuntimeError("Simulate.backward: the Adjoint ran another forward after this output was "
                                   "computed; use one Adjoint per differentiable call")
            d4 = np.zeros((adj.n, 4), np.float32)
            d4[:, :3] = grad_pos.detach().cpu().numpy()
            g_v0, _, g_g, g_mu = adj.backward(d4)
            return (torch.from_numpy(g_v0[:, :3].copy()), torch.tensor(g_mu, dtype=torch.float64),
                    torch.from_numpy(g_g.astype(np.float64)), None, None, None, None, None, None)

    return Simulate


def soft_histogram(pos, half, bins=32, xp=np):
    """The density histogram of the loss: each particle's unit mass spread
    over the 8 nearest of bins^3 bin centres of the box [-half, half] with
    trilinear weights (piecewise linear in the positions, so differentiable).
    `xp` is numpy or torch."""
    half = xp.asarray(half, dtype=pos.dtype) if xp is np else xp.tensor(half, dtype=pos.dtype)
    u = (pos + half) / (2.0 * half) * bins - 0.5
    u = xp.clip(u, 0.0, bins - 1.0 - 1e-6) if xp is np else u.clamp(0.0, bins - 1.0 - 1e-6)
    i0 = xp.floor(u)
    f = u - i0
    i0 = i0.astype(np.int64) if xp is np else i0.long()
    H = xp.zeros(bins ** 3, dtype=pos.dtype)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((f[:, 0] if dx else 1.0 - f[:, 0]) * (f[:, 1] if dy else 1.0 - f[:, 1]) *
                     (f[:, 2] if dz else 1.0 - f[:, 2]))
                idx = ((i0[:, 2] + dz) * bins + (i0[:, 1] + dy)) * bins + (i0[:, 0] + dx)
                if xp is np:
                    np.add.at(H, idx, w)
                else:
                    H = H.index_add(0, idx, w)
    return H


def loss(pos, half, c_target, h_target, bins=32, xp=np):
    """|centroid - c*|^2 + |H - H*|^2 over the final positions (N, 3)."""
    c = pos.mean(0)
    H = soft_histogram(pos, half, bins, xp)
    ct = xp.asarray(c_target, dtype=pos.dtype) if xp is np else xp.tensor(c_target, dtype=pos.dtype)
    return ((c - ct) ** 2).sum() + ((H - h_target) ** 2).sum()
