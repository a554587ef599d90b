"""MI355X-native DFSPH step: Python plumbing over libfluid_sim.so.

The step itself is hand-written HIP for gfx950 behind the C ABI in
include/fluid_sim.h; this package loads that library (and raises if it is not
built), packs the reference's parameter blocks through it, and builds the
seeded benchmark scenes.
"""
from ._lib import FluidError, LIB_PATH, build, lib
from .sim import (FIELDS, HIST_BINS, KERNELS, NBR_CAP, Batch, Grid, Sim, body_force, device_count, fluid_hash, half_bits, kernel,
                  lattice_dims, pack_sim_params, pack_step, pack_tracer, refine_passes, scan_u32, seed_slab,
                  seed_tracers, spacing_for, substeps_for)
from . import scenes

__all__ = [
    "FluidError", "LIB_PATH", "build", "lib", "FIELDS", "KERNELS", "NBR_CAP", "Batch", "Grid", "Sim", "body_force",
    "device_count", "fluid_hash", "half_bits", "kernel", "lattice_dims", "pack_sim_params", "pack_step",
    "pack_tracer", "refine_passes", "scan_u32", "seed_slab", "seed_tracers", "spacing_for", "substeps_for",
    "scenes",
]
