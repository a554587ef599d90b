#ifndef FLUID_FFI_H
#define FLUID_FFI_H

/*
 * The reference's renderer C ABI (crates/fluid-ffi/include/fluid_ffi.h:1-188),
 * same symbols and struct layouts, served by the MI355X DFSPH step in
 * headless mode: `metal_layer` may be NULL, frames simulate and do not
 * present. Declarations keep the reference's order and types; the
 * comments say where behaviour differs.
 */

#include <stdarg.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>

/* The renderer behind the C ABI: an opaque box, driven from one thread. */
typedef struct FluidRenderer FluidRenderer;

typedef struct FluidVec3 {  /* fluid_ffi.h:17-21 */
    float x;
    float y;
    float z;
} FluidVec3;

typedef struct FluidRenderStats {  /* fluid_ffi.h:23-47 */
    uint64_t frames;
    float interval_p50_us;
    float interval_p99_us;
    float interval_max_us;
    float acquire_p50_us;
    float acquire_p99_us;
    float encode_p50_us;
    float encode_p99_us;
    float gpu_p50_us;
    float gpu_p99_us;
    float compression_avg;
    float compression_max;
    float density_min;
    float density_max;
    float pressure_min;
    float pressure_max;
    float v_max;
    float temperature_min;
    float temperature_max;
    uint32_t clamp_count;
    uint32_t neighbour_overflow;
    uint32_t substeps;
    uint64_t idle_frames;
} FluidRenderStats;

#ifdef __cplusplus
extern "C" {
#endif

/* How many fingers the sim drags with at once (fluid_ffi.h:58). */
extern const uint32_t FLUID_TOUCH_SLOTS;

/*
 * fluid_ffi.h:74-83. A nonzero `sim_substeps` builds the DFSPH fluid at
 * `particle_scale` times the shipped count (or at `bench_spacing` metres when
 * it is nonzero), capped at `sim_substeps` substeps a frame, with `tracers`
 * tracers, on the current HIP device. The sprite demo (sim_substeps = 0) and
 * the stage-0 microbench (bench_sweeps > 0) are not part of the step: create
 * returns NULL for them, with the reason on stderr, as it does when setup
 * fails. `metal_layer` is not read and may be NULL.
 */
struct FluidRenderer *fluid_renderer_create(void *metal_layer,
                                            uint32_t width,
                                            uint32_t height,
                                            uint32_t particle_count,
                                            float sprite_radius,
                                            uint32_t bench_sweeps,
                                            float bench_spacing,
                                            uint32_t sim_substeps,
                                            uint32_t tracers,
                                            float particle_scale);

/*
 * fluid_ffi.h:95-99: one frame of the sim (render.rs:3021-3494 without the
 * draw). Returns 1 when it stepped, 0 when the settled sim slept.
 */
uint32_t fluid_renderer_frame(struct FluidRenderer *renderer,
                              struct FluidVec3 gravity,
                              struct FluidVec3 user_acceleration,
                              struct FluidVec3 rotation_rate,
                              double now_ms);

/* fluid_ffi.h:110: reseed at `scale` times the shipped count. */
void fluid_renderer_set_particles(struct FluidRenderer *renderer, float scale);

/* fluid_ffi.h:120: the count set_particles(scale) would seed. */
uint32_t fluid_renderer_particles_at(const struct FluidRenderer *renderer, float scale);

/* fluid_ffi.h:132-136: one finger, normalised drawable coordinates. */
void fluid_renderer_touch(struct FluidRenderer *renderer,
                          uint32_t slot,
                          float x,
                          float y,
                          bool down);

/*
 * fluid_ffi.h:155-161: the look is recorded (headless mode draws nothing);
 * it decides, as in the reference, whether tracers advect (the glass look
 * only) and it restarts the idle gate.
 */
void fluid_renderer_set_look(struct FluidRenderer *renderer,
                             bool flat,
                             bool particles,
                             struct FluidVec3 low,
                             bool gradient,
                             struct FluidVec3 high,
                             uint32_t lens);

/* fluid_ffi.h:168: recorded; the box extent is fixed at create, as in the reference's sim. */
void fluid_renderer_resize(struct FluidRenderer *renderer, uint32_t width, uint32_t height);

/* fluid_ffi.h:175: stats by value; acquire and encode timings read zero headless. */
struct FluidRenderStats fluid_renderer_stats(const struct FluidRenderer *renderer);

/* fluid_ffi.h:182 */
void fluid_renderer_destroy(struct FluidRenderer *renderer);

#ifdef __cplusplus
}  /* extern "C" */
#endif

#endif /* FLUID_FFI_H */
