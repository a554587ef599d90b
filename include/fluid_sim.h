#ifndef FLUID_SIM_H
#define FLUID_SIM_H

/*
 * Step-level C ABI of the MI355X DFSPH step.
 *
 * The reference keeps its step crate-private behind `Sim` (render.rs:1382-2226)
 * and drives it with two byte blocks: the 64-byte SimParams uniform
 * (pack_sim_params, crates/fluid-core/src/sim.rs:325-350) and the 144-byte
 * Step immediates (pack_step, sim.rs:200-252), dispatching a substep as
 * count, scan, scatter, density_div, div_apply, forces_eval,
 * forces_den_apply, refine x (den_kappa, den_apply), integrate, then per
 * frame reduce_stats and the tracer kernels (render.rs:3168-3228). This
 * header exposes exactly that surface, taking the two blocks verbatim, so a
 * Rust (or any C) caller of the reference's step binds it unchanged.
 *
 * Conventions (the reference's fluid_ffi.h ones): an opaque handle; NULL or
 * a negative status on failure with the reason on stderr and in
 * fluid_sim_last_error(); one thread per handle. Kernels run on the handle's
 * HIP stream; calls that return host data synchronise it.
 *
 * Every function has no counterpart in fluid_ffi.h unless its comment names
 * one; each cites the reference code whose role it takes.
 */

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct FluidSim FluidSim;

#define FLUID_SIM_PARAMS_BYTES 64u  /* sim.rs:330 */
#define FLUID_SIM_STEP_BYTES 144u   /* sim.rs:202 STEP_BYTES */
#define FLUID_SIM_STATS 14u         /* sim_solve.wgsl:61 stats[14] */
#define FLUID_SIM_NBR_CAP 96u       /* sim_solve.wgsl:388 NBR_CAP */

#define FLUID_OK 0
#define FLUID_ERR_ARG (-1)
#define FLUID_ERR_HIP (-2)
#define FLUID_ERR_NOMEM (-3)
#define FLUID_ERR_DEVICE (-4)

/* Fields for fluid_sim_download_field, all in slot (cell) order. */
#define FLUID_FIELD_DENSITY 0u    /* f32 [N]    sim_solve.wgsl:501 */
#define FLUID_FIELD_ALPHA 1u      /* f32 [N]    :509 */
#define FLUID_FIELD_KD 2u         /* f32 [N]    kappa/rho, :516, 589, 695 */
#define FLUID_FIELD_PRESSURE 3u   /* f32 [N]    :590, 696 */
#define FLUID_FIELD_RHO_NEAR 4u   /* f32 [N]    positions.w during a substep, :507 */
#define FLUID_FIELD_ACCEL 5u      /* f32 [N][4] :577 */
#define FLUID_FIELD_XSPH 6u       /* f32 [N][4] :581 */
#define FLUID_FIELD_WALL_GRAD 7u  /* f32 [N][4] :499 */
#define FLUID_FIELD_WORK_POS 8u   /* f32 [N][4] working positions */
#define FLUID_FIELD_WORK_VEL 9u   /* f32 [N][4] working velocities */

/* Message of the last failure on this thread ("" if none). */
const char *fluid_sim_last_error(void);

/* Number of HIP devices visible (0 without a GPU). */
int fluid_sim_device_count(void);

/*
 * Builds a sim on the current HIP device from a 64-byte SimParams block
 * (count at byte 28 is the particle count). Takes the role of Sim::new
 * (render.rs:1468-1612): velocities and previous velocities start at zero,
 * temperature at 293.15 K, previous pressure at zero; positions start at
 * zero until fluid_sim_upload_state. `tracers` packed tracers are allocated
 * (zero = none). `nbr_cap` is the neighbour-list cap (0 = NBR_CAP, 96; a multiple of 4, at most 256).
 * Returns NULL on failure.
 */
FluidSim *fluid_sim_create(const uint8_t *params64, uint32_t tracers, uint32_t nbr_cap);

/* Frees every device buffer; the handle is dead afterwards. */
void fluid_sim_destroy(FluidSim *sim);

uint32_t fluid_sim_count(const FluidSim *sim);
uint32_t fluid_sim_cells(const FluidSim *sim);
uint32_t fluid_sim_tracer_count(const FluidSim *sim);

/*
 * Resting-set upload (Appendix of the record layouts: positions vec4 with
 * w = 0, velocities vec4 with w = acceleration EMA, prev_vel vec4 with
 * w = pressure EMA, prev_pressure f32, temperature f32). Any pointer may be
 * NULL to leave that record unchanged. Host pointers; synchronous.
 */
int fluid_sim_upload_state(FluidSim *sim, const float *pos4, const float *vel4, const float *prev_vel4,
                           const float *prev_pressure, const float *temperature);
int fluid_sim_download_state(FluidSim *sim, float *pos4, float *vel4, float *prev_vel4, float *prev_pressure,
                             float *temperature);

/* Scratch field of the last substep (FLUID_FIELD_*), 1 or 4 floats a particle. */
int fluid_sim_download_field(FluidSim *sim, uint32_t field, float *out);

/*
 * Grid of the last rebuild: cell id per resting particle (N), per-cell
 * counts (C), starts (C + 1, starts[C] = N), stable slot -> resting-index
 * permutation (N), neighbour counts (N). Any pointer may be NULL.
 */
int fluid_sim_download_grid(FluidSim *sim, uint32_t *cellid, uint32_t *counts, uint32_t *starts, uint32_t *perm,
                            uint32_t *nbr_n);

/* Neighbour list of the last density sweep: N x cap slot indices. */
int fluid_sim_download_neighbours(FluidSim *sim, uint32_t *nbr);

/*
 * Carries a particle identity (initially the upload index) through every
 * sort, so callers can follow particles across substeps. Off by default: it
 * adds 8 bytes a particle to the scatter.
 */
int fluid_sim_track_ids(FluidSim *sim, bool enable);
int fluid_sim_download_ids(FluidSim *sim, uint32_t *ids);

/*
 * `n` substeps of the reference's dispatch order (render.rs:3175-3203) with
 * one Step block (dt in it is the substep length), `refine` density
 * iterations each (render.rs:77-79: 5 or 2). dt = 0 encodes no solve, as
 * render.rs:3165-3167 does. Asynchronous.
 */
int fluid_sim_substeps(FluidSim *sim, const uint8_t *step144, uint32_t n, uint32_t refine);

/*
 * Tolerance mode of one sim, the batch's fluid_batch_set_tolerance and
 * fluid_batch_set_div_tolerance for a single scene: the density solve
 * iterates to `tol` (min_passes .. max_passes <= 256 passes; the substeps'
 * `refine` is then ignored) and the divergence solve to `div_tol` (at most
 * div_max_passes <= 256 passes). A tolerance <= 0 keeps that solve's
 * reference schedule; both <= 0 is the reference step. In this mode
 * fluid_sim_substeps waits on the device between passes: it stops issuing a
 * solve's passes once that solve has converged (a few passes stay queued
 * ahead of the wait, and those past convergence do nothing), so it returns
 * while the last passes and integrate are still running.
 */
int fluid_sim_set_tolerance(FluidSim *sim, float tol, uint32_t min_passes, uint32_t max_passes, float div_tol,
                            uint32_t div_max_passes);
/* The solve's tile size (64, 128 or 256 slots a workgroup; 128 at
 * creation). Results do not depend on it. */
int fluid_sim_set_tile(FluidSim *sim, uint32_t tile);
uint32_t fluid_sim_tile(const FluidSim *sim);
/* Passes of the last substep: out[0] density, out[1] divergence. */
int fluid_sim_iterations(FluidSim *sim, uint32_t out[2]);
/* The pass-count histograms of every substep under tolerance mode since the
 * last reset, as fluid_batch_pass_histogram: 2 x 257 words. */
int fluid_sim_pass_histogram(FluidSim *sim, uint32_t *out, bool reset);

/*
 * Grid rebuild plus the density sweep alone: the zero-substep path of the
 * reference's test driver (render.rs:3902-3937). Asynchronous.
 */
int fluid_sim_density_only(FluidSim *sim, const uint8_t *step144);

/* reduce_stats (sim_solve.wgsl:800-874) into the device stats block. */
int fluid_sim_reduce_stats(FluidSim *sim);
/* Copies the 14-float stats block (synchronises). Layout: sim_solve.wgsl:819-873. */
int fluid_sim_read_stats(FluidSim *sim, float *out14);

/* splat, resolve, advect (sim_tracers.wgsl:98-183) with a frame Step block. */
int fluid_sim_tracers(FluidSim *sim, const uint8_t *step144);
int fluid_sim_upload_tracers(FluidSim *sim, const uint32_t *packed2);
int fluid_sim_download_tracers(FluidSim *sim, uint32_t *packed2);

/*
 * One frame as render.rs:3164-3228 encodes it: `n` substeps of `step144`,
 * reduce_stats, then the tracers with `tracer_step144` (when the sim has
 * tracers and tracer_step144 is not NULL).
 */
int fluid_sim_frame(FluidSim *sim, const uint8_t *step144, uint32_t n, uint32_t refine,
                    const uint8_t *tracer_step144);

/* The sim's 64-byte SimParams block (sim.rs:330-350). */
int fluid_sim_params(const FluidSim *sim, uint8_t *out64);

/*
 * The resting set on the device, for consumers that read it in place as the
 * reference's sprite, body-splat and disc pullers bind it
 * (render.rs:1818-1828): positions (vec4, w = 0), velocities (vec4, w =
 * acceleration EMA), prev_vel (vec4, w = pressure EMA), prev_pressure,
 * temperature, the last substep's density in the same slot order, and the
 * packed tracers (2 words each; NULL without tracers). The pointers stay
 * valid for the sim's life (a reseed makes a new sim). `stream` is the
 * hipStream_t the sim's kernels run on: work queued on it after a frame
 * reads that frame's state, or a caller records an event on it to wait for.
 * The sim rewrites these arrays on its next substep.
 */
typedef struct FluidDeviceRecords {
    float *positions;
    float *velocities;
    float *prev_vel;
    float *prev_pressure;
    float *temperature;
    float *density;
    uint32_t *tracers;
    uint32_t count;
    uint32_t tracer_count;
    void *stream;
} FluidDeviceRecords;
int fluid_sim_device_records(FluidSim *sim, FluidDeviceRecords *out);

/* Cumulative CFL clamps and neighbour-list overflow (stats[9], stats[10]). */
int fluid_sim_counters(FluidSim *sim, uint64_t *clamps, uint64_t *overflow);

/* Waits for the handle's stream. */
int fluid_sim_synchronize(FluidSim *sim);

/*
 * Per-kernel HIP-event timing of the launches that follow, on the handle's
 * stream (the timing adds an event pair around each timed launch).
 * `kernel_mask` selects kernel classes by bit (FLUID_KERNEL_*); 0 stops.
 */
#define FLUID_KERNEL_GRID_COUNT 0u
#define FLUID_KERNEL_GRID_SCAN 1u
#define FLUID_KERNEL_GRID_SCATTER 2u
#define FLUID_KERNEL_DENSITY_DIV 3u
#define FLUID_KERNEL_DIV_APPLY 4u
/* forces_eval and forces_den_apply (sim_solve.wgsl:524-663) run as one
 * sweep, timed as FLUID_KERNEL_FORCES; class 6 has no launches. */
#define FLUID_KERNEL_FORCES 5u
#define FLUID_KERNEL_UNUSED 6u
#define FLUID_KERNEL_DEN_KAPPA 7u
#define FLUID_KERNEL_DEN_APPLY 8u
#define FLUID_KERNEL_INTEGRATE 9u
#define FLUID_KERNEL_REDUCE_STATS 10u
#define FLUID_KERNEL_TRACERS 11u
#define FLUID_KERNEL_CLASSES 12u
int fluid_sim_profile(FluidSim *sim, uint32_t kernel_mask);
/* The same, timing one launch in every `stride` of each class in the mask
 * (stride 0 or 1: every launch). A timed launch's start and stop events go
 * into the stream beside it and slow it a little (about 2% of a 1M substep
 * with every den_kappa launch timed); the read-out's averages cover the
 * timed launches. No counterpart in the reference. */
int fluid_sim_profile_sampled(FluidSim *sim, uint32_t kernel_mask, uint32_t stride);
/* Total milliseconds and launch count per kernel class since profiling began (synchronises). */
int fluid_sim_profile_read(FluidSim *sim, double *ms, uint64_t *launches);
const char *fluid_sim_kernel_name(uint32_t kernel);

/* A JSON description of the handle's device and launch configuration. */
int fluid_sim_describe(const FluidSim *sim, char *buf, size_t size);

/*
 * The grid's exclusive scan on its own (the decoupled-look-back scan that
 * replaces scan_single, sim_scan.wgsl:97-120): starts[c] = sum of
 * counts[0..c), starts[cells] = total. Host arrays; synchronous.
 */
int fluid_scan_u32(const uint32_t *counts, uint32_t cells, uint32_t *starts);

/*
 * The same scan without any allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). `counts`
 * (cells words) and `starts` (cells + 1 words) are DEVICE pointers on the
 * current device; `stream` is a hipStream_t (NULL: the null stream). The call
 * is asynchronous on that stream, leaves `counts` unchanged and writes
 * starts[c] = sum of counts[0..c), starts[cells] = total. At most 2^24 cells
 * (FLUID_ERR_ARG beyond, nothing launched). Its look-back state lives in a
 * 64 KB device array that each GPU the library runs on holds until the
 * process ends, shared by every call on that GPU: calls that overlap on one
 * GPU race and must be ordered by the caller (one stream, or events).
 */
int fluid_scan_u32_noalloc(const uint32_t *counts, uint32_t cells, uint32_t *starts, void *stream);

/* ---- Host helpers: the reference's CPU substrate and schedule ------------
 * Pure host functions (no device needed), C++ restatements of
 * crates/fluid-core/src/sim.rs:1-350, lib.rs:34-36 and render.rs:57-106,
 * 1362-1380. The grid takes three half-extents: the reference's Grid::new
 * takes two and fixes z at SLAB_DEPTH / 2 (sim.rs:24-26).
 */
typedef struct FluidGrid {
    float min[3];
    float cell;
    uint32_t dims[3];
} FluidGrid;

uint32_t fluid_hash(uint32_t x);                                                  /* sim.rs:318-323 */
void fluid_grid_new(const float half[3], float support, FluidGrid *out);          /* sim.rs:24-32 */
uint32_t fluid_grid_cell_of(const FluidGrid *grid, const float pos[3]);           /* sim.rs:38-43 */
float fluid_kernel(float r, float h);                                             /* sim.rs:49-60 */
void fluid_pack_sim_params(const FluidGrid *grid, uint32_t count, float h, float mass, float rho0,
                           uint8_t *out64);                                       /* sim.rs:330-350 */
/* touches20: 5 fingers x (x, y, vx, vy), may be NULL. */
void fluid_pack_step(const float force[3], const float omega[3], const float domega[3], float dt, float v_clamp,
                     uint32_t seed, const float *touches20, uint32_t touch_n, float touch_r,
                     uint8_t *out144);                                            /* sim.rs:215-252 */
void fluid_body_force(const float gravity[3], const float user_acceleration[3], float out[3]); /* lib.rs:34-36 */
void fluid_lattice_dims(float spacing, const float half[3], float fill, uint32_t out[3]);     /* sim.rs:65-72 */
/* Writes 4 floats a particle when out4 is not NULL; returns the count. */
uint32_t fluid_seed_slab(float spacing, const float half[3], float fill, float *out4);       /* sim.rs:77-100 */
void fluid_seed_tracers(uint32_t count, const float half[3], float fill, uint32_t *out2);    /* sim.rs:105-122 */
void fluid_pack_tracer(const float pos[3], float speed, const float extent[2], uint32_t out[2]); /* sim.rs:127-133 */
uint16_t fluid_half_bits(float x);                                                /* sim.rs:139-151 */
uint32_t fluid_refine_passes(float dt_sub);                                       /* render.rs:77-79 */
uint32_t fluid_substeps_for(float dt, float v_max, float spacing, uint32_t cap);  /* render.rs:97-106 */
float fluid_spacing_for(float scale, const float extent[2]);                      /* render.rs:1362-1380 */
float fluid_slab_depth(void);                                                     /* sim.rs:10 */

/* The reference's compile-time physical coefficients as runtime values,
 * phys5 = {mu, k_near, tension_scale, adhesion_scale, xsph_rate}: dynamic
 * viscosity (DYNAMIC_VISCOSITY, sim_solve.wgsl:135), the near-pressure
 * stiffness (K_NEAR, :137), factors on tension_gamma and adhesion_beta
 * (:161-172) and the XSPH rate (:657). Defaults {1.002e-3, 3000, 1, 1, 6}
 * are the reference's; the adjoint's differentiable profile is
 * {mu, 0, 0, 0, 0}. No counterpart in the reference (compile-time there). */
int fluid_sim_set_physics(FluidSim *sim, const float *phys5);

/* ---- the frame driver behind fluid_renderer_* -------------------------------
 *
 * Additions to the reference's renderer ABI (include/fluid_ffi.h), for
 * callers that drive or inspect what fluid_renderer_frame does. The renderer
 * handle is fluid_ffi.h's FluidRenderer (named by its struct tag here).
 */

/* The renderer's sim (NULL in the stage-0 microbench mode): its resting set
 * is what the reference's sprite and disc pullers bind (render.rs:1818-1828);
 * see fluid_sim_device_records. Owned by the renderer. */
FluidSim *fluid_renderer_sim(struct FluidRenderer *r);
/* The stage-0 microbench mode (fluid_renderer_create with bench_sweeps > 0,
 * render.rs:2429-2729): the first frame's verdicts (validate_bench: the scan
 * against a CPU scan, every 16th density against a CPU brute force to 1e-3;
 * 1 pass, 0 fail, -1 before the first frame), the particle and cell counts,
 * and, when the pointers are not NULL, the last frame's densities (seed
 * order), the starts (cells + 1) and the seeded positions (4 floats a
 * particle). FLUID_ERR_ARG outside the bench mode. */
int fluid_renderer_bench(struct FluidRenderer *r, int *scan_ok, int *density_ok, uint32_t *count, uint32_t *cells,
                         float *density, uint32_t *starts, float *pos4);
/* The last stepped frame's substep Step block, substep count, density passes
 * and tracer Step block (render.rs:3169-3228): replaying them on another sim
 * from the same state gives the same state. */
int fluid_renderer_last_step(const struct FluidRenderer *r, uint8_t *step144, uint32_t *substeps, uint32_t *refine,
                             uint8_t *tracer_step144);
/* The live ramp of a flat or particle look painted with a gradient
 * (Ramp::follow, render.rs:708-743): 1 and its two ends, or 0 when no ramp
 * is live (a solid or glass look, or no stats yet). */
int fluid_renderer_ramp(const struct FluidRenderer *r, float *ends2);
/* Lens::ends and Lens::floor (render.rs:649-680) of lens code `lens` (0
 * velocity, 1 acceleration, 2 pressure, 3 proximity, 4 direction; other
 * codes paint velocity, lib.rs:215-221) from a 14-float stats block. */
int fluid_lens_ends(uint32_t lens, const float *stats14, float spacing, float *ends2, float *floor);

/* The input filters of the frame driver, as the renderer runs them each
 * frame: ForceFilter (render.rs:862-903), RotationTracker (986-1015),
 * Fingers (912-984) and IdleGate (1024-1104). Pure host code. */
typedef struct FluidFilters FluidFilters;
FluidFilters *fluid_filters_create(void);
void fluid_filters_destroy(FluidFilters *f);
/* ForceFilter::apply: writes the smoothed body force, returns the deviation. */
float fluid_filters_force(FluidFilters *f, const float body_force[3], float smooth[3]);
/* RotationTracker::apply: the smoothed rate and its derivative. */
void fluid_filters_rotation(FluidFilters *f, const float rate[3], float dt, float omega[3], float domega[3]);
/* A finger at drawable point (x, y) in [0, 1]^2 (top left 0, 0), or lifted. */
void fluid_filters_touch(FluidFilters *f, uint32_t slot, float x, float y, bool down);
/* Fingers::step: the live fingers packed forward as (x, y, vx, vy) in box
 * metres, their count returned, their shared radius in *radius. */
uint32_t fluid_filters_fingers(FluidFilters *f, const float extent[2], float dt, float *touches20, float *radius);
/* IdleGate::asleep. */
bool fluid_filters_asleep(FluidFilters *f, bool touched, const float smooth[3], float dev, const float omega[3],
                          float v_max);
/* IdleGate::STILL_FRAMES: still frames before the gate sleeps. */
uint32_t fluid_filters_still_frames(void);

/* ---- batched step ---------------------------------------------------------
 *
 * Many independent boxes of uneven particle counts stepped together, one
 * launch per kernel per substep. The reference steps one box per Sim
 * (render.rs:1382-1612, 3168-3203); a batch is the same step over a list of
 * SimParams blocks, each scene with its own Step and density-pass count. A
 * scene's results are those of a lone FluidSim on the same inputs, bit for
 * bit. Scenes are indexed 0..scenes-1; state and fields are per scene, in
 * the scene's slot order. Same conventions as FluidSim.
 */
typedef struct FluidBatch FluidBatch;

/* `scenes` SimParams blocks of 64 bytes each, back to back (count at byte 28
 * of each is that scene's particle count). NULL on failure. */
FluidBatch *fluid_batch_create(uint32_t scenes, const uint8_t *params64, uint32_t nbr_cap);
void fluid_batch_destroy(FluidBatch *batch);
uint32_t fluid_batch_scene_count(const FluidBatch *batch);
/* Slots of the shared particle set (each scene padded to 256 slots). */
uint32_t fluid_batch_slot_count(const FluidBatch *batch);
int fluid_batch_scene_info(const FluidBatch *batch, uint32_t scene, uint32_t *slot_begin, uint32_t *count,
                           uint32_t *cells);
/* Per scene, as fluid_sim_upload_state / fluid_sim_download_state. */
int fluid_batch_upload_state(FluidBatch *batch, uint32_t scene, const float *pos4, const float *vel4,
                             const float *prev_vel4, const float *prev_pressure, const float *temperature);
int fluid_batch_download_state(FluidBatch *batch, uint32_t scene, float *pos4, float *vel4, float *prev_vel4,
                               float *prev_pressure, float *temperature);
int fluid_batch_download_field(FluidBatch *batch, uint32_t scene, uint32_t field, float *out);
/* The scene's grid in its own numbering (cell ids from 0, starts from 0). */
int fluid_batch_download_grid(FluidBatch *batch, uint32_t scene, uint32_t *cellid, uint32_t *counts,
                              uint32_t *starts, uint32_t *perm);
/* One substep of every scene: `steps144` holds one 144-byte Step per scene;
 * `refine` one density-pass count per scene (render.rs:77-79). A scene
 * whose Step has dt <= 0 sits the substep out (render.rs:3165-3167). */
int fluid_batch_substep(FluidBatch *batch, const uint8_t *steps144, const uint32_t *refine);
/* Tolerance mode of the density solve (tol > 0): each scene runs density
 * passes until the mean of max(rho* - rho0, 0) / rho0 over its particles is
 * <= tol, at least min_passes and at most max_passes (<= 256); `refine` is
 * then ignored. A pass that meets the tolerance still applies its
 * correction and is the last. tol <= 0 restores the reference schedule.
 * The reference fixes the pass count (sim_solve.wgsl:674-703, render.rs:
 * 77-79); this iterates the same den_kappa / den_apply pair to a tolerance,
 * as DFSPH's density solver does. */
int fluid_batch_set_tolerance(FluidBatch *batch, float tol, uint32_t min_passes, uint32_t max_passes);
/* Tolerance mode of the divergence solve (div_tol > 0): after density_div's
 * kappa and div_apply (the reference's one pass, sim_solve.wgsl:508-516,
 * 665-668), each scene runs further passes of (divergence kappa from the
 * current velocities, div_apply) while the mean of max(Drho/Dt, 0) dt / rho0
 * of the last pass is above div_tol, at most max_passes (<= 256) in all.
 * div_tol <= 0 restores the one pass. */
int fluid_batch_set_div_tolerance(FluidBatch *batch, float div_tol, uint32_t max_passes);
/* Whether tolerance mode runs the passes of each solve under tolerance on
 * work lists (the default: each pass's sweeps take only the tiles of the
 * scenes still iterating) or dispatches every tile each pass. The results are the
 * same bits. On work lists a substep is issued directly, never from a
 * graph: fluid_batch_substep waits on the device between passes and stops
 * issuing a solve's passes once every scene has converged (a few passes
 * stay queued ahead of the wait, and those past convergence do nothing). */
int fluid_batch_persistent_passes(FluidBatch *batch, bool enable);
/* Density passes each scene ran in the last substep under tolerance mode. */
int fluid_batch_iterations(FluidBatch *batch, uint32_t *out);
/* Divergence passes each scene ran in the last substep under tolerance mode. */
int fluid_batch_div_iterations(FluidBatch *batch, uint32_t *out);
/* The pass-count histograms of every substep under tolerance mode since the
 * last reset, 2 x 257 words: out[p] counts the (scene, substep) pairs that ran
 * p density passes, out[257 + p] those that ran p divergence passes. `out`
 * may be NULL; `reset` clears them afterwards. */
int fluid_batch_pass_histogram(FluidBatch *batch, uint32_t *out, bool reset);
/* Capture each substep shape once in a hipGraph and replay it (except
 * tolerance mode on work lists, issued directly to stop early). */
int fluid_batch_use_graph(FluidBatch *batch, bool enable);
/* Per scene, as fluid_sim_reduce_stats / fluid_sim_read_stats: out holds
 * scenes x 14 floats. */
int fluid_batch_reduce_stats(FluidBatch *batch);
int fluid_batch_read_stats(FluidBatch *batch, float *out);
/* scenes x (clamps, overflow). */
int fluid_batch_counters(FluidBatch *batch, uint64_t *out);
int fluid_batch_synchronize(FluidBatch *batch);
int fluid_batch_profile(FluidBatch *batch, uint32_t kernel_mask);
int fluid_batch_profile_read(FluidBatch *batch, double *ms, uint64_t *launches);
/* The batch's launch shape: the solve's tile size (64, 128 or 256 slots a
 * workgroup), the LDS staging capacity (slots; a tile whose stencil rows
 * need more gathers from global memory), the scene-group size (slots; 0 runs
 * the sweeps over the whole batch at once) and the streams the groups go
 * round (1 to 4). None of them changes a result bit. */
typedef struct FluidBatchTuning {
    uint32_t tile;
    uint32_t stage_cap;
    uint32_t group_slots;
    uint32_t streams;
} FluidBatchTuning;
int fluid_batch_get_tuning(const FluidBatch *batch, FluidBatchTuning *out);
/* Sets a launch shape (stage_cap is rounded down to 64s and capped at the
 * tile size's largest). */
int fluid_batch_set_tuning(FluidBatch *batch, const FluidBatchTuning *tuning);
/* The autotuner: times `trials` substeps of each candidate launch shape from
 * a saved copy of the state (restored afterwards, the cumulative counters
 * included) and keeps the fastest, choosing in turn the tile size, the
 * staging capacity, then the scene groups and streams. Writes the chosen
 * capacity to *chosen; fluid_batch_get_tuning reads the whole shape. */
int fluid_batch_autotune(FluidBatch *batch, const uint8_t *steps144, const uint32_t *refine, uint32_t trials,
                         uint32_t *chosen);
int fluid_batch_describe(const FluidBatch *batch, char *buf, size_t size);
/* As fluid_sim_set_physics, for every scene of the batch. */
int fluid_batch_set_physics(FluidBatch *batch, const float *phys5);

/* ---- adjoint ----------------------------------------------------------------
 *
 * The gradient of a loss on the final state of `steps` substeps with respect
 * to the initial velocities and positions, the dynamic viscosity and each
 * substep's gravity. The reference has no adjoint; this is the hand-written
 * reverse of the whole step above: viscosity, near pressure, cohesion, wall
 * adhesion, the XSPH blend, the rotating frame and both pressure solves, at
 * the settings of fluid_adjoint_set_physics ({mu, k_near, tension_scale,
 * adhesion_scale, xsph_rate}; the default is the reference's step). Steps
 * must have no touches (the forward refuses them: the finger pull has no
 * reverse). The grid permutation and neighbour lists are held fixed per
 * substep. The
 * forward keeps a checkpoint of the resting set every `checkpoint_every`
 * substeps; the backward replays each segment with a tape and runs the
 * reverse sweeps (the density solve's passes as transposed Jacobi sweeps).
 */
typedef struct FluidAdjoint FluidAdjoint;

FluidAdjoint *fluid_adjoint_create(const uint8_t *params64, uint32_t steps, uint32_t checkpoint_every,
                                   uint32_t nbr_cap);
void fluid_adjoint_destroy(FluidAdjoint *adj);
int fluid_adjoint_set_physics(FluidAdjoint *adj, const float *phys5);
/* Runs the substeps from positions pos4 and velocities vel4 (NULL: zero),
 * previous pressures zero, temperature 293.15 K. steps144 holds one Step per
 * substep (dt > 0), refine one density-pass count per substep (<= 8). Writes
 * the final resting positions and velocities (NULL: skipped); their order is
 * the final resting order. Fails if a neighbour list overflowed its cap. */
int fluid_adjoint_forward(FluidAdjoint *adj, const float *pos4, const float *vel4, const uint8_t *steps144,
                          const uint32_t *refine, float *pos_out4, float *vel_out4);
/* Given d loss / d final position (and velocity; NULL: zero), in the order
 * fluid_adjoint_forward returned them, writes d loss / d initial velocity and
 * position (input order; NULL: skipped), d loss / d gravity of each substep
 * (steps x 3) and d loss / d mu. */
int fluid_adjoint_backward(FluidAdjoint *adj, const float *dpos4, const float *dvel4, float *dvel0_4,
                           float *dpos0_4, float *dgravity3, float *dmu);
/* GPU time of the last forward and backward (HIP events on the stream). */
int fluid_adjoint_timing(const FluidAdjoint *adj, double *forward_ms, double *backward_ms);

#ifdef __cplusplus
}
#endif

#endif /* FLUID_SIM_H */
