"""Loads libfluid_sim.so (the HIP DFSPH step and its C ABI) and declares the
ctypes signatures of include/fluid_sim.h.

There is no fallback: when the library is missing, loading raises. The
library is built in-tree by the package Makefile (see build()).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libfluid_sim.so")
HEADERS = [os.path.join(os.path.dirname(ROOT), "include", h) for h in ("fluid_sim.h", "fluid_ffi.h")]

_lib = None


def build(jobs: int = 8) -> str:
    """Compiles every HIP source of the package for gfx950 (make in the package dir)."""
    subprocess.run(["make", "-B", f"-j{jobs}", "-C", ROOT], check=True)
    return LIB_PATH


def _declare(L):
    vp, u32, f32, f64, i32, u64 = C.c_void_p, C.c_uint32, C.c_float, C.c_double, C.c_int, C.c_uint64
    P = C.POINTER
    u8p = P(C.c_uint8)
    sig = {
        "fluid_sim_last_error": ([], C.c_char_p),
        "fluid_sim_device_count": ([], i32),
        "fluid_sim_create": ([u8p, u32, u32], vp),
        "fluid_sim_destroy": ([vp], None),
        "fluid_sim_count": ([vp], u32),
        "fluid_sim_cells": ([vp], u32),
        "fluid_sim_tracer_count": ([vp], u32),
        "fluid_sim_upload_state": ([vp] + [P(f32)] * 5, i32),
        "fluid_sim_download_state": ([vp] + [P(f32)] * 5, i32),
        "fluid_sim_download_field": ([vp, u32, P(f32)], i32),
        "fluid_sim_download_grid": ([vp] + [P(u32)] * 5, i32),
        "fluid_sim_download_neighbours": ([vp, P(u32)], i32),
        "fluid_sim_track_ids": ([vp, C.c_bool], i32),
        "fluid_sim_download_ids": ([vp, P(u32)], i32),
        "fluid_sim_substeps": ([vp, u8p, u32, u32], i32),
        "fluid_sim_set_tolerance": ([vp, f32, u32, u32, f32, u32], i32),
        "fluid_sim_set_tile": ([vp, u32], i32),
        "fluid_sim_tile": ([vp], u32),
        "fluid_sim_iterations": ([vp, P(u32)], i32),
        "fluid_sim_pass_histogram": ([vp, P(u32), C.c_bool], i32),
        "fluid_sim_density_only": ([vp, u8p], i32),
        "fluid_sim_reduce_stats": ([vp], i32),
        "fluid_sim_read_stats": ([vp, P(f32)], i32),
        "fluid_sim_tracers": ([vp, u8p], i32),
        "fluid_sim_upload_tracers": ([vp, P(u32)], i32),
        "fluid_sim_download_tracers": ([vp, P(u32)], i32),
        "fluid_sim_frame": ([vp, u8p, u32, u32, u8p], i32),
        "fluid_sim_counters": ([vp, P(u64), P(u64)], i32),
        "fluid_sim_synchronize": ([vp], i32),
        "fluid_sim_profile": ([vp, u32], i32),
        "fluid_sim_profile_sampled": ([vp, u32, u32], i32),
        "fluid_sim_profile_read": ([vp, P(f64), P(u64)], i32),
        "fluid_sim_kernel_name": ([u32], C.c_char_p),
        "fluid_sim_describe": ([vp, C.c_char_p, C.c_size_t], i32),
        "fluid_scan_u32": ([P(u32), u32, P(u32)], i32),
        "fluid_scan_u32_noalloc": ([vp, u32, vp, vp], i32),
        "fluid_sim_set_physics": ([vp, P(f32)], i32),
        "fluid_batch_set_physics": ([vp, P(f32)], i32),
        "fluid_adjoint_create": ([u8p, u32, u32, u32], vp),
        "fluid_adjoint_destroy": ([vp], None),
        "fluid_adjoint_set_physics": ([vp, P(f32)], i32),
        "fluid_adjoint_forward": ([vp, P(f32), P(f32), u8p, P(u32), P(f32), P(f32)], i32),
        "fluid_adjoint_backward": ([vp, P(f32), P(f32), P(f32), P(f32), P(f32), P(f32)], i32),
        "fluid_adjoint_timing": ([vp, P(f64), P(f64)], i32),
        "fluid_batch_create": ([u32, u8p, u32], vp),
        "fluid_batch_destroy": ([vp], None),
        "fluid_batch_scene_count": ([vp], u32),
        "fluid_batch_slot_count": ([vp], u32),
        "fluid_batch_scene_info": ([vp, u32, P(u32), P(u32), P(u32)], i32),
        "fluid_batch_upload_state": ([vp, u32] + [P(f32)] * 5, i32),
        "fluid_batch_download_state": ([vp, u32] + [P(f32)] * 5, i32),
        "fluid_batch_download_field": ([vp, u32, u32, P(f32)], i32),
        "fluid_batch_download_grid": ([vp, u32] + [P(u32)] * 4, i32),
        "fluid_batch_substep": ([vp, u8p, P(u32)], i32),
        "fluid_batch_set_tolerance": ([vp, f32, u32, u32], i32),
        "fluid_batch_iterations": ([vp, P(u32)], i32),
        "fluid_batch_set_div_tolerance": ([vp, f32, u32], i32),
        "fluid_batch_get_tuning": ([vp, P(u32)], i32),
        "fluid_batch_set_tuning": ([vp, P(u32)], i32),
        "fluid_batch_div_iterations": ([vp, P(u32)], i32),
        "fluid_batch_persistent_passes": ([vp, C.c_bool], i32),
        "fluid_batch_pass_histogram": ([vp, P(u32), C.c_bool], i32),
        "fluid_batch_use_graph": ([vp, C.c_bool], i32),
        "fluid_batch_reduce_stats": ([vp], i32),
        "fluid_batch_read_stats": ([vp, P(f32)], i32),
        "fluid_batch_counters": ([vp, P(u64)], i32),
        "fluid_batch_synchronize": ([vp], i32),
        "fluid_batch_profile": ([vp, u32], i32),
        "fluid_batch_profile_read": ([vp, P(f64), P(u64)], i32),
        "fluid_batch_autotune": ([vp, u8p, P(u32), u32, P(u32)], i32),
        "fluid_batch_describe": ([vp, C.c_char_p, C.c_size_t], i32),
        "fluid_hash": ([u32], u32),
        "fluid_grid_new": ([P(f32), f32, vp], None),
        "fluid_grid_cell_of": ([vp, P(f32)], u32),
        "fluid_kernel": ([f32, f32], f32),
        "fluid_pack_sim_params": ([vp, u32, f32, f32, f32, u8p], None),
        "fluid_pack_step": ([P(f32), P(f32), P(f32), f32, f32, u32, P(f32), u32, f32, u8p], None),
        "fluid_body_force": ([P(f32), P(f32), P(f32)], None),
        "fluid_lattice_dims": ([f32, P(f32), f32, P(u32)], None),
        "fluid_seed_slab": ([f32, P(f32), f32, P(f32)], u32),
        "fluid_seed_tracers": ([u32, P(f32), f32, P(u32)], None),
        "fluid_pack_tracer": ([P(f32), f32, P(f32), P(u32)], None),
        "fluid_half_bits": ([f32], C.c_uint16),
        "fluid_refine_passes": ([f32], u32),
        "fluid_substeps_for": ([f32, f32, f32, u32], u32),
        "fluid_spacing_for": ([f32, P(f32)], f32),
        "fluid_slab_depth": ([], f32),
        "fluid_renderer_sim": ([vp], vp),
        "fluid_sim_params": ([vp, u8p], i32),
        "fluid_sim_device_records": ([vp, vp], i32),
        "fluid_renderer_bench": ([vp, P(i32), P(i32), P(u32), P(u32), P(f32), P(u32), P(f32)], i32),
        "fluid_renderer_last_step": ([vp, u8p, P(u32), P(u32), u8p], i32),
        "fluid_renderer_ramp": ([vp, P(f32)], i32),
        "fluid_lens_ends": ([u32, P(f32), f32, P(f32), P(f32)], i32),
        "fluid_filters_create": ([], vp),
        "fluid_filters_destroy": ([vp], None),
        "fluid_filters_force": ([vp, P(f32), P(f32)], f32),
        "fluid_filters_rotation": ([vp, P(f32), f32, P(f32), P(f32)], None),
        "fluid_filters_touch": ([vp, u32, f32, f32, C.c_bool], None),
        "fluid_filters_fingers": ([vp, P(f32), f32, P(f32), P(f32)], u32),
        "fluid_filters_asleep": ([vp, C.c_bool, P(f32), f32, P(f32), f32], C.c_bool),
        "fluid_filters_still_frames": ([], u32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def lib():
    """The loaded library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run __graft_entry__.build() (make -C {ROOT}); "
                "the DFSPH step has no other implementation")
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


class FluidError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().fluid_sim_last_error().decode(errors="replace")
        raise FluidError(f"{what} failed ({rc}): {msg}")
