"""Python view of the C ABI (include/fluid_sim.h): the step's parameter
packers and a `Sim` handle that mirrors the reference's Sim
(crates/fluid-core/src/render.rs:1382-2226) and its dispatch order
(render.rs:3164-3228). Plumbing for tests and the bench; every computation
runs in libfluid_sim.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

FIELDS = {
    "density": (0, 1),
    "alpha": (1, 1),
    "kd": (2, 1),
    "pressure": (3, 1),
    "rho_near": (4, 1),
    "accel": (5, 4),
    "xsph": (6, 4),
    "wall_grad": (7, 4),
    "work_pos": (8, 4),
    "work_vel": (9, 4),
}
# Bins of the tolerance-mode pass histograms: pass counts 0 .. 256.
HIST_BINS = 257

# Kernel classes of the timer, by FLUID_KERNEL_* index ("forces" is
# forces_eval and forces_den_apply in one sweep; "unused" has no launches).
KERNELS = ["grid_count", "grid_scan", "grid_scatter", "density_div", "div_apply", "forces",
           "unused", "den_kappa", "den_apply", "integrate", "reduce_stats", "tracers"]
NBR_CAP = 96
STATS = 14


class _GridC(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("cell", C.c_float), ("dims", C.c_uint32 * 3)]


def _f32p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _u32p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def _u8p(b: bytes):
    arr = (C.c_uint8 * len(b)).from_buffer_copy(bytes(b))
    return arr


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


@dataclass
class Grid:
    """sim.rs:14-44 with three half-extents."""
    min: tuple
    cell: float
    dims: tuple

    @staticmethod
    def new(half, support) -> "Grid":
        g = _GridC()
        lib().fluid_grid_new(_f3(half), float(support), C.byref(g))
        return Grid(tuple(g.min), float(g.cell), tuple(int(d) for d in g.dims))

    def _c(self):
        g = _GridC()
        g.min = (C.c_float * 3)(*self.min)
        g.cell = self.cell
        g.dims = (C.c_uint32 * 3)(*self.dims)
        return g

    def cell_count(self) -> int:
        return int(self.dims[0]) * int(self.dims[1]) * int(self.dims[2])

    def cell_of(self, pos) -> int:
        g = self._c()
        return int(lib().fluid_grid_cell_of(C.byref(g), _f3(pos)))


def pack_sim_params(grid: Grid, count: int, h: float, mass: float, rho0: float) -> bytes:
    out = (C.c_uint8 * 64)()
    g = grid._c()
    lib().fluid_pack_sim_params(C.byref(g), int(count), float(h), float(mass), float(rho0), out)
    return bytes(out)


def pack_step(force=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), domega=(0.0, 0.0, 0.0), dt=0.0, v_clamp=0.0,
              seed=0, touches=None, touch_n=0, touch_r=0.0) -> bytes:
    """sim.rs:215-252. `touches`: up to 5 rows of (x, y, vx, vy)."""
    out = (C.c_uint8 * 144)()
    t = np.zeros((5, 4), np.float32)
    if touches is not None:
        tt = np.asarray(touches, np.float32).reshape(-1, 4)
        t[: len(tt)] = tt
    lib().fluid_pack_step(_f3(force), _f3(omega), _f3(domega), float(dt), float(v_clamp), int(seed) & 0xFFFFFFFF,
                          _f32p(t), int(touch_n), float(touch_r), out)
    return bytes(out)


def kernel(r: float, h: float) -> float:
    return float(lib().fluid_kernel(float(r), float(h)))


def lattice_dims(spacing, half, fill):
    out = np.zeros(3, np.uint32)
    lib().fluid_lattice_dims(float(spacing), _f3(half), float(fill), _u32p(out))
    return tuple(int(x) for x in out)


def seed_slab(spacing, half, fill) -> np.ndarray:
    n = lib().fluid_seed_slab(float(spacing), _f3(half), float(fill), None)
    out = np.zeros((n, 4), np.float32)
    lib().fluid_seed_slab(float(spacing), _f3(half), float(fill), _f32p(out))
    return out


def seed_tracers(count, half, fill) -> np.ndarray:
    out = np.zeros((count, 2), np.uint32)
    lib().fluid_seed_tracers(int(count), _f3(half), float(fill), _u32p(out))
    return out


def pack_tracer(pos, speed, extent):
    out = np.zeros(2, np.uint32)
    lib().fluid_pack_tracer(_f3(pos), float(speed), (C.c_float * 2)(*extent), _u32p(out))
    return out


def half_bits(x: float) -> int:
    return int(lib().fluid_half_bits(float(x)))


def fluid_hash(x: int) -> int:
    return int(lib().fluid_hash(int(x) & 0xFFFFFFFF))


def refine_passes(dt_sub: float) -> int:
    return int(lib().fluid_refine_passes(float(dt_sub)))


def substeps_for(dt: float, v_max: float, spacing: float, cap: int) -> int:
    return int(lib().fluid_substeps_for(float(dt), float(v_max), float(spacing), int(cap)))


def spacing_for(scale: float, extent) -> float:
    return float(lib().fluid_spacing_for(float(scale), (C.c_float * 2)(*extent)))


def body_force(gravity, user_acceleration):
    out = (C.c_float * 3)()
    lib().fluid_body_force(_f3(gravity), _f3(user_acceleration), out)
    return tuple(out)


def device_count() -> int:
    return int(lib().fluid_sim_device_count())


def scan_u32(counts: np.ndarray) -> np.ndarray:
    c = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros(c.size + 1, np.uint32)
    check(lib().fluid_scan_u32(_u32p(c), c.size, _u32p(out)), "fluid_scan_u32")
    return out


class Sim:
    """One scene on the GPU behind the C ABI."""

    def __init__(self, params64: bytes, tracers: int = 0, nbr_cap: int = 0):
        assert len(params64) == 64
        self.L = lib()
        self.h = self.L.fluid_sim_create(_u8p(params64), int(tracers), int(nbr_cap))
        if not self.h:
            raise RuntimeError("fluid_sim_create failed: " + self.L.fluid_sim_last_error().decode())
        self.n = int(self.L.fluid_sim_count(self.h))
        self.cells = int(self.L.fluid_sim_cells(self.h))
        self.ntr = int(self.L.fluid_sim_tracer_count(self.h))
        self.cap = int(nbr_cap) if nbr_cap else NBR_CAP

    def close(self):
        if getattr(self, "h", None):
            self.L.fluid_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_state(self, pos4=None, vel4=None, prev_vel4=None, prev_pressure=None, temperature=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32)
                for a in (pos4, vel4, prev_vel4, prev_pressure, temperature)]
        check(self.L.fluid_sim_upload_state(self.h, *[_f32p(a) for a in arrs]), "upload_state")

    def state(self) -> dict:
        n = self.n
        out = dict(pos=np.empty((n, 4), np.float32), vel=np.empty((n, 4), np.float32),
                   prev_vel=np.empty((n, 4), np.float32), prev_pressure=np.empty(n, np.float32),
                   temperature=np.empty(n, np.float32))
        check(self.L.fluid_sim_download_state(self.h, *[_f32p(out[k]) for k in
                                                        ("pos", "vel", "prev_vel", "prev_pressure",
                                                         "temperature")]), "download_state")
        return out

    def field(self, name: str) -> np.ndarray:
        which, per = FIELDS[name]
        out = np.empty(self.n * per, np.float32)
        check(self.L.fluid_sim_download_field(self.h, which, _f32p(out)), "download_field")
        return out.reshape(self.n, per) if per > 1 else out

    def grid(self) -> dict:
        n, c = self.n, self.cells
        out = dict(cellid=np.empty(n, np.uint32), counts=np.empty(c, np.uint32),
                   starts=np.empty(c + 1, np.uint32), perm=np.empty(n, np.uint32),
                   nbr_n=np.empty(n, np.uint32))
        check(self.L.fluid_sim_download_grid(self.h, *[_u32p(out[k]) for k in
                                                       ("cellid", "counts", "starts", "perm", "nbr_n")]),
              "download_grid")
        return out

    def neighbours(self) -> np.ndarray:
        out = np.empty((self.n, self.cap), np.uint32)
        check(self.L.fluid_sim_download_neighbours(self.h, _u32p(out)), "download_neighbours")
        return out

    def track_ids(self, enable=True):
        check(self.L.fluid_sim_track_ids(self.h, bool(enable)), "track_ids")

    def ids(self) -> np.ndarray:
        out = np.empty(self.n, np.uint32)
        check(self.L.fluid_sim_download_ids(self.h, _u32p(out)), "download_ids")
        return out

    def substeps(self, step144: bytes, n: int, refine: int):
        check(self.L.fluid_sim_substeps(self.h, _u8p(step144), int(n), int(refine)), "substeps")

    def set_tolerance(self, tol: float = 0.0, min_passes: int = 1, max_passes: int = 0, div_tol: float = 0.0,
                      div_max_passes: int = 0):
        """Tolerance mode: the density solve iterates until its mean error is
        at most `tol` (min_passes .. max_passes passes), the divergence solve
        until its mean error is at most `div_tol` (at most div_max_passes);
        a tolerance of 0 keeps that solve's reference schedule."""
        check(self.L.fluid_sim_set_tolerance(self.h, float(tol), int(min_passes), int(max_passes), float(div_tol),
                                             int(div_max_passes)), "sim set_tolerance")

    def set_tile(self, tile: int):
        """The solve's tile size: 64, 128 or 256 slots a workgroup."""
        check(self.L.fluid_sim_set_tile(self.h, int(tile)), "sim set_tile")

    @property
    def tile(self) -> int:
        return int(self.L.fluid_sim_tile(self.h))

    def iterations(self) -> tuple:
        """(density passes, divergence passes) of the last substep."""
        out = np.zeros(2, np.uint32)
        check(self.L.fluid_sim_iterations(self.h, _u32p(out)), "sim iterations")
        return int(out[0]), int(out[1])

    def pass_histogram(self, reset: bool = False) -> np.ndarray:
        """[2, HIST_BINS] counts of substeps by passes run (density, divergence)."""
        out = np.zeros(2 * HIST_BINS, np.uint32)
        check(self.L.fluid_sim_pass_histogram(self.h, _u32p(out), bool(reset)), "sim pass_histogram")
        return out.reshape(2, HIST_BINS)

    def density_only(self, step144: bytes):
        check(self.L.fluid_sim_density_only(self.h, _u8p(step144)), "density_only")

    def reduce_stats(self):
        check(self.L.fluid_sim_reduce_stats(self.h), "reduce_stats")

    def read_stats(self) -> np.ndarray:
        out = np.empty(STATS, np.float32)
        check(self.L.fluid_sim_read_stats(self.h, _f32p(out)), "read_stats")
        return out

    def tracers(self, step144: bytes):
        check(self.L.fluid_sim_tracers(self.h, _u8p(step144)), "tracers")

    def upload_tracers(self, packed: np.ndarray):
        a = np.ascontiguousarray(packed, np.uint32)
        assert a.size == 2 * self.ntr
        check(self.L.fluid_sim_upload_tracers(self.h, _u32p(a)), "upload_tracers")

    def download_tracers(self) -> np.ndarray:
        out = np.empty((self.ntr, 2), np.uint32)
        check(self.L.fluid_sim_download_tracers(self.h, _u32p(out)), "download_tracers")
        return out

    def frame(self, step144: bytes, n: int, refine: int, tracer_step144: bytes | None = None):
        ts = _u8p(tracer_step144) if tracer_step144 is not None else None
        check(self.L.fluid_sim_frame(self.h, _u8p(step144), int(n), int(refine), ts), "frame")

    def counters(self):
        a, b = C.c_uint64(), C.c_uint64()
        check(self.L.fluid_sim_counters(self.h, C.byref(a), C.byref(b)), "counters")
        return int(a.value), int(b.value)

    def synchronize(self):
        check(self.L.fluid_sim_synchronize(self.h), "synchronize")

    def set_physics(self, mu=1.002e-3, k_near=3000.0, tension_scale=1.0, adhesion_scale=1.0, xsph_rate=6.0):
        arr = (C.c_float * 5)(mu, k_near, tension_scale, adhesion_scale, xsph_rate)
        check(self.L.fluid_sim_set_physics(self.h, arr), "set_physics")

    def describe(self) -> dict:
        import json
        buf = C.create_string_buffer(512)
        check(self.L.fluid_sim_describe(self.h, buf, 512), "describe")
        return json.loads(buf.value.decode())

    def profile(self, kernels=None, stride: int = 1):
        """Times the launches of the named kernel classes (one in every
        `stride` of each class)."""
        mask = 0
        for k in (kernels or []):
            mask |= 1 << (KERNELS.index(k) if isinstance(k, str) else int(k))
        check(self.L.fluid_sim_profile_sampled(self.h, mask, int(stride)), "profile")

    def profile_read(self) -> dict:
        ms = np.zeros(len(KERNELS), np.float64)
        cnt = np.zeros(len(KERNELS), np.uint64)
        check(self.L.fluid_sim_profile_read(self.h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                            cnt.ctypes.data_as(C.POINTER(C.c_uint64))), "profile_read")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNELS) if cnt[i]}


class Batch:
    """Many scenes stepped together behind the batched C ABI
    (include/fluid_sim.h, fluid_batch_*). Scene s's state, fields and grid
    read as a lone Sim's would."""

    def __init__(self, params: list, nbr_cap: int = 0):
        blob = b"".join(params)
        assert len(blob) == 64 * len(params) and params
        self.L = lib()
        self.h = self.L.fluid_batch_create(len(params), _u8p(blob), int(nbr_cap))
        if not self.h:
            raise RuntimeError("fluid_batch_create failed: " + self.L.fluid_sim_last_error().decode())
        self.scenes = int(self.L.fluid_batch_scene_count(self.h))
        self.slots = int(self.L.fluid_batch_slot_count(self.h))
        self.info = []
        for s in range(self.scenes):
            a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
            check(self.L.fluid_batch_scene_info(self.h, s, C.byref(a), C.byref(b), C.byref(c)), "scene_info")
            self.info.append((int(a.value), int(b.value), int(c.value)))

    def count(self, s: int) -> int:
        return self.info[s][1]

    def close(self):
        if getattr(self, "h", None):
            self.L.fluid_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload_state(self, s, pos4=None, vel4=None, prev_vel4=None, prev_pressure=None, temperature=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32)
                for a in (pos4, vel4, prev_vel4, prev_pressure, temperature)]
        check(self.L.fluid_batch_upload_state(self.h, int(s), *[_f32p(a) for a in arrs]), "batch upload_state")

    def state(self, s) -> dict:
        n = self.count(s)
        out = dict(pos=np.empty((n, 4), np.float32), vel=np.empty((n, 4), np.float32),
                   prev_vel=np.empty((n, 4), np.float32), prev_pressure=np.empty(n, np.float32),
                   temperature=np.empty(n, np.float32))
        check(self.L.fluid_batch_download_state(self.h, int(s), *[_f32p(out[k]) for k in
                                                                  ("pos", "vel", "prev_vel", "prev_pressure",
                                                                   "temperature")]), "batch download_state")
        return out

    def field(self, s, name: str) -> np.ndarray:
        which, per = FIELDS[name]
        n = self.count(s)
        out = np.empty(n * per, np.float32)
        check(self.L.fluid_batch_download_field(self.h, int(s), which, _f32p(out)), "batch download_field")
        return out.reshape(n, per) if per > 1 else out

    def grid(self, s) -> dict:
        n, c = self.count(s), self.info[s][2]
        out = dict(cellid=np.empty(n, np.uint32), counts=np.empty(c, np.uint32),
                   starts=np.empty(c + 1, np.uint32), perm=np.empty(n, np.uint32))
        check(self.L.fluid_batch_download_grid(self.h, int(s), *[_u32p(out[k]) for k in
                                                                 ("cellid", "counts", "starts", "perm")]),
              "batch download_grid")
        return out

    def substep(self, steps144: list, refine):
        blob = b"".join(steps144)
        assert len(blob) == 144 * self.scenes
        r = np.ascontiguousarray(refine, np.uint32)
        assert r.size == self.scenes
        check(self.L.fluid_batch_substep(self.h, _u8p(blob), _u32p(r)), "batch substep")

    def set_tolerance(self, tol: float, min_passes: int = 1, max_passes: int = 0):
        check(self.L.fluid_batch_set_tolerance(self.h, float(tol), int(min_passes), int(max_passes)),
              "batch set_tolerance")

    def set_div_tolerance(self, div_tol: float, max_passes: int = 0):
        check(self.L.fluid_batch_set_div_tolerance(self.h, float(div_tol), int(max_passes)),
              "batch set_div_tolerance")

    def iterations(self) -> np.ndarray:
        out = np.empty(self.scenes, np.uint32)
        check(self.L.fluid_batch_iterations(self.h, _u32p(out)), "batch iterations")
        return out

    def tuning(self) -> dict:
        """The launch shape: tile, stage_cap, group_slots, streams."""
        out = np.zeros(4, np.uint32)
        check(self.L.fluid_batch_get_tuning(self.h, _u32p(out)), "batch get_tuning")
        return dict(zip(("tile", "stage_cap", "group_slots", "streams"), map(int, out)))

    def set_tuning(self, tile: int, stage_cap: int, group_slots: int, streams: int):
        arr = np.array([tile, stage_cap, group_slots, streams], np.uint32)
        check(self.L.fluid_batch_set_tuning(self.h, _u32p(arr)), "batch set_tuning")

    def persistent_passes(self, enable=True):
        """Tolerance mode's passes on work lists of the tiles still iterating
        (default) or over every tile; the same bits either way."""
        check(self.L.fluid_batch_persistent_passes(self.h, bool(enable)), "batch persistent_passes")

    def div_iterations(self) -> np.ndarray:
        out = np.empty(self.scenes, np.uint32)
        check(self.L.fluid_batch_div_iterations(self.h, _u32p(out)), "batch div_iterations")
        return out

    def pass_histogram(self, reset: bool = False) -> np.ndarray:
        """[2, HIST_BINS] counts of (scene, substep) pairs by passes run (density,
        divergence) over the substeps under tolerance mode."""
        out = np.zeros(2 * HIST_BINS, np.uint32)
        check(self.L.fluid_batch_pass_histogram(self.h, _u32p(out), bool(reset)), "batch pass_histogram")
        return out.reshape(2, HIST_BINS)

    def use_graph(self, enable=True):
        check(self.L.fluid_batch_use_graph(self.h, bool(enable)), "batch use_graph")

    def reduce_stats(self):
        check(self.L.fluid_batch_reduce_stats(self.h), "batch reduce_stats")

    def read_stats(self) -> np.ndarray:
        out = np.empty((self.scenes, STATS), np.float32)
        check(self.L.fluid_batch_read_stats(self.h, _f32p(out)), "batch read_stats")
        return out

    def counters(self) -> np.ndarray:
        out = np.empty((self.scenes, 2), np.uint64)
        check(self.L.fluid_batch_counters(self.h, out.ctypes.data_as(C.POINTER(C.c_uint64))), "batch counters")
        return out

    def synchronize(self):
        check(self.L.fluid_batch_synchronize(self.h), "batch synchronize")

    def profile(self, kernels=None):
        mask = 0
        for k in (kernels or []):
            mask |= 1 << (KERNELS.index(k) if isinstance(k, str) else int(k))
        check(self.L.fluid_batch_profile(self.h, mask), "batch profile")

    def profile_read(self) -> dict:
        ms = np.zeros(len(KERNELS), np.float64)
        cnt = np.zeros(len(KERNELS), np.uint64)
        check(self.L.fluid_batch_profile_read(self.h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                              cnt.ctypes.data_as(C.POINTER(C.c_uint64))), "batch profile_read")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNELS) if cnt[i]}

    def autotune(self, steps144: list, refine, trials: int = 3) -> int:
        blob = b"".join(steps144)
        r = np.ascontiguousarray(refine, np.uint32)
        out = C.c_uint32()
        check(self.L.fluid_batch_autotune(self.h, _u8p(blob), _u32p(r), int(trials), C.byref(out)),
              "batch autotune")
        return int(out.value)

    def set_physics(self, mu=1.002e-3, k_near=3000.0, tension_scale=1.0, adhesion_scale=1.0, xsph_rate=6.0):
        arr = (C.c_float * 5)(mu, k_near, tension_scale, adhesion_scale, xsph_rate)
        check(self.L.fluid_batch_set_physics(self.h, arr), "batch set_physics")

    def describe(self) -> dict:
        import json
        buf = C.create_string_buffer(512)
        check(self.L.fluid_batch_describe(self.h, buf, 512), "batch describe")
        return json.loads(buf.value.decode())
