// The reference's renderer C ABI (crates/fluid-ffi/src/lib.rs:71-285) served
// headless by the MI355X step: the frame driver of render.rs (input filters
// 862-1105, the frame's CPU schedule 3021-3099, the encode 3164-3228, stats
// 3545-3575) in C++ over include/fluid_sim.h. Nothing is drawn.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/fluid_ffi.h"
#include "../../include/fluid_sim.h"
#include "fluid_kernels.hpp"

namespace {

constexpr float WORLD_SCALE = 4.0f;                                    // lib.rs:13
constexpr float METRES_PER_PIXEL = WORLD_SCALE * 0.0254f / 458.0f;     // render.rs:19
constexpr float MAX_DT = 1.0f / 30.0f;                                 // render.rs:23
constexpr float SIM_SPACING = WORLD_SCALE * 0.0025f;                   // render.rs:1353
constexpr float REST_DENSITY = 998.2f;                                 // sim.rs:6
constexpr int MAX_TOUCHES = 5;                                         // sim.rs:180
constexpr int RING = 240;                                              // render.rs:14

struct Ring {  // render.rs:108-146
    std::array<float, RING> samples{};
    size_t filled = 0, next = 0;
    void push(float v) {
        samples[next] = v;
        next = (next + 1) % RING;
        filled = std::min<size_t>(filled + 1, RING);
    }
    float percentile(float p) const {
        if (filled == 0) return 0.0f;
        std::vector<float> s(samples.begin(), samples.begin() + (long)filled);
        std::sort(s.begin(), s.end());
        return s[(size_t)std::lround(p * (float)(filled - 1))];
    }
    float max() const {
        float m = 0.0f;
        for (size_t i = 0; i < filled; ++i) m = std::max(m, samples[i]);
        return m;
    }
};

struct ForceFilter {  // render.rs:862-903
    float smooth[3] = {0, 0, 0};
    bool started = false;
    float apply(const float force[3], float out[3]) {
        if (!started) {
            started = true;
            std::memcpy(smooth, force, sizeof smooth);
        }
        float dev = 0.0f;
        for (int i = 0; i < 3; ++i) dev += (force[i] - smooth[i]) * (force[i] - smooth[i]);
        dev = std::sqrt(dev);
        const float alpha = std::min(std::max(dev / 2.0f, 0.02f), 1.0f);
        for (int i = 0; i < 3; ++i) smooth[i] += (force[i] - smooth[i]) * alpha;
        std::memcpy(out, smooth, sizeof smooth);
        return dev;
    }
};

struct RotationTracker {  // render.rs:986-1015
    float smooth[3] = {0, 0, 0};
    bool primed = false;
    void apply(const float omega[3], float dt, float out[3], float domega[3]) {
        if (!primed || dt <= 0.0f || dt > 0.1f) {
            primed = true;
            std::memcpy(smooth, omega, sizeof smooth);
            std::memcpy(out, omega, sizeof smooth);
            domega[0] = domega[1] = domega[2] = 0.0f;
            return;
        }
        const float alpha = 1.0f - std::exp(-dt / 0.025f);
        float prev[3];
        std::memcpy(prev, smooth, sizeof prev);
        for (int i = 0; i < 3; ++i) smooth[i] += (omega[i] - smooth[i]) * alpha;
        for (int i = 0; i < 3; ++i) domega[i] = std::min(std::max((smooth[i] - prev[i]) / dt, -200.0f), 200.0f);
        std::memcpy(out, smooth, sizeof smooth);
    }
};

struct Touches {
    float each[MAX_TOUCHES][4] = {};
    uint32_t count = 0;
    float radius = 0.0f;
};

struct Fingers {  // render.rs:912-984
    static constexpr float RADIUS = WORLD_SCALE * 0.025f;
    static constexpr float SMOOTH_TAU = 0.02f;
    bool down[MAX_TOUCHES] = {};
    float at[MAX_TOUCHES][2] = {};
    float world[MAX_TOUCHES][2] = {};
    float velocity[MAX_TOUCHES][2] = {};
    bool tracked[MAX_TOUCHES] = {};
    bool any_down() const {
        for (bool d : down)
            if (d) return true;
        return false;
    }
    Touches step(const float extent[2], float dt) {
        Touches t;
        t.radius = RADIUS;
        for (int slot = 0; slot < MAX_TOUCHES; ++slot) {
            if (!down[slot]) {
                tracked[slot] = false;
                continue;
            }
            const float w[2] = {(2.0f * at[slot][0] - 1.0f) * extent[0], (1.0f - 2.0f * at[slot][1]) * extent[1]};
            if (tracked[slot] && dt > 0.0f) {
                const float alpha = 1.0f - std::exp(-dt / SMOOTH_TAU);
                for (int a = 0; a < 2; ++a)
                    velocity[slot][a] += ((w[a] - world[slot][a]) / dt - velocity[slot][a]) * alpha;
            } else {
                velocity[slot][0] = velocity[slot][1] = 0.0f;
            }
            world[slot][0] = w[0];
            world[slot][1] = w[1];
            tracked[slot] = true;
            float* e = t.each[t.count];
            e[0] = w[0];
            e[1] = w[1];
            e[2] = velocity[slot][0];
            e[3] = velocity[slot][1];
            t.count++;
        }
        return t;
    }
};

struct IdleGate {  // render.rs:1024-1104
    static constexpr float V_SLEEP = 0.04f, DEV_SLEEP = 0.5f, DEV_WAKE = 1.2f;
    static constexpr uint32_t STILL_FRAMES = 180;
    static constexpr float COS_WAKE = 0.999657f, MAG_WAKE = 0.3f, OMEGA_WAKE = 0.05f;
    uint32_t still = 0;
    bool has_sleep = false;
    float sleep_force[3] = {0, 0, 0};
    bool sleeping() const { return has_sleep; }
    static float dot(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
    bool asleep(bool touched, const float smooth[3], float dev, const float omega[3], float v_max) {
        if (touched) {
            has_sleep = false;
            still = 0;
            return false;
        }
        const bool spinning = std::sqrt(dot(omega, omega)) > OMEGA_WAKE;
        if (has_sleep) {
            const float cos = dot(smooth, sleep_force) /
                              std::max(std::sqrt(dot(smooth, smooth) * dot(sleep_force, sleep_force)), 1e-6f);
            const float mag_shift = std::fabs(std::sqrt(dot(smooth, smooth)) - std::sqrt(dot(sleep_force, sleep_force)));
            if (dev > DEV_WAKE || cos < COS_WAKE || mag_shift > MAG_WAKE || spinning) {
                has_sleep = false;
                still = 0;
                return false;
            }
            return true;
        }
        if (v_max < V_SLEEP && dev < DEV_SLEEP && !spinning) {
            still += 1;
            if (still >= STILL_FRAMES) {
                has_sleep = true;
                std::memcpy(sleep_force, smooth, sizeof sleep_force);
                return true;
            }
        } else {
            still = 0;
        }
        return false;
    }
};

// The lens a ramp paints with (render.rs:608-680), by the code
// fluid_renderer_set_look takes (lib.rs:211-222; 0 and unknown codes are
// velocity): the two ends the frame's stats block sets, and the narrowest
// span worth a ramp, in the lens's units.
constexpr float STANDARD_GRAVITY = 9.80665f;  // lib.rs:5
void lens_ends(uint32_t lens, const float* stats, float out[2]) {
    switch (lens) {
        case 1: out[0] = stats[12]; out[1] = stats[13]; break;  // acceleration
        case 2: out[0] = stats[4]; out[1] = stats[5]; break;    // pressure
        case 3: out[0] = stats[2]; out[1] = stats[3]; break;    // proximity (density)
        default: out[0] = stats[11]; out[1] = stats[6]; break;  // velocity, direction
    }
}
float lens_floor(uint32_t lens, float spacing) {
    switch (lens) {
        case 1: return STANDARD_GRAVITY;
        case 2: return REST_DENSITY * STANDARD_GRAVITY * spacing;
        case 3: return 0.01f * REST_DENSITY;
        default: return std::sqrt(2.0f * STANDARD_GRAVITY * spacing);
    }
}

// The live ramp's ends, chased frame by frame (render.rs:682-743): opening
// out with tau 0.15 s, closing in with 0.6 s, never narrower than the
// lens's floor. None (false) until a stats block with density has landed.
struct Ramp {
    static constexpr float OPEN_TAU = 0.15f, CLOSE_TAU = 0.6f;
    bool has_lens = false;
    uint32_t lens = 0;
    float ends[2] = {0.0f, 0.0f};
    bool follow(uint32_t l, const float* stats, float spacing, float dt, float out[2]) {
        if (stats[3] <= 0.0f) return false;
        const float floor = lens_floor(l, spacing);
        float want[2];
        lens_ends(l, stats, want);
        want[1] = std::max(want[1], want[0] + floor);
        if (!has_lens || lens != l) {
            has_lens = true;
            lens = l;
            ends[0] = want[0];
            ends[1] = want[1];
        } else {
            for (int i = 0; i < 2; ++i) {
                const bool opening = i == 0 ? want[i] < ends[i] : want[i] > ends[i];
                const float tau = opening ? OPEN_TAU : CLOSE_TAU;
                ends[i] += (want[i] - ends[i]) * (1.0f - std::exp(-dt / tau));
            }
            ends[1] = std::max(ends[1], ends[0] + floor);
        }
        out[0] = ends[0];
        out[1] = ends[1];
        return true;
    }
};

// Sim::new (render.rs:1468-1612) on the C ABI: the seeded slab, tracers and
// parameter block at `spacing` in a box of half-extent `extent` and slab depth.
FluidSim* build_sim(const float extent[2], float spacing, uint32_t tracers, uint32_t* count_out) {
    const float half[3] = {extent[0], extent[1], fluid_slab_depth() * 0.5f};
    const float h = 1.2f * spacing;
    FluidGrid grid;
    fluid_grid_new(half, 2.0f * h, &grid);
    const uint32_t count = fluid_seed_slab(spacing, half, 0.5f, nullptr);
    std::vector<float> pos((size_t)count * 4);
    fluid_seed_slab(spacing, half, 0.5f, pos.data());
    uint8_t params[64];
    fluid_pack_sim_params(&grid, count, h, REST_DENSITY * spacing * spacing * spacing, REST_DENSITY, params);
    FluidSim* sim = fluid_sim_create(params, tracers, 0);
    if (!sim) return nullptr;
    if (fluid_sim_upload_state(sim, pos.data(), nullptr, nullptr, nullptr, nullptr) != FLUID_OK) {
        fluid_sim_destroy(sim);
        return nullptr;
    }
    if (tracers > 0) {
        std::vector<uint32_t> tr((size_t)tracers * 2);
        fluid_seed_tracers(tracers, half, 0.5f, tr.data());
        fluid_sim_upload_tracers(sim, tr.data());
    }
    fprintf(stderr, "sim: %u particles, %u cells, spacing %g m\n", count, fluid_sim_cells(sim), spacing);
    *count_out = count;
    return sim;
}

}  // namespace

// The frame driver's four input filters (render.rs:862-1105), one object so
// that the renderer and the C ABI's fluid_filters_* share them.
struct FluidFilters {
    ForceFilter force;
    RotationTracker rotation;
    Fingers fingers;
    IdleGate idle;
};

struct FluidRenderer {
    FluidSim* sim = nullptr;
    // The stage-0 microbench mode (bench_sweeps > 0) in place of the sim.
    fluid::Stage0Bench* bench = nullptr;
    float extent[2] = {0, 0};
    float spacing = 0.0f;
    uint32_t max_substeps = 0, tracers = 0, count = 0, width = 0, height = 0;
    uint32_t substeps_used = 0, frame_seed = 0;
    uint64_t frames = 0, idle_frames = 0;
    double last_frame_ms = 0.0;
    float stats[FLUID_SIM_STATS] = {};
    bool glass = true;  // the reference's default look, Look::Glass
    // A flat or particle look painting a ramp with `lens` (set_look).
    bool ramp_look = false;
    uint32_t lens = 0;
    Ramp ramp;
    bool ramp_live = false;
    float ramp_ends[2] = {0.0f, 0.0f};
    // The last stepped frame's Step blocks and schedule (for callers that
    // replay or inspect it).
    uint8_t last_step[144] = {}, last_tracer_step[144] = {};
    uint32_t last_refine = 0;
    FluidFilters filters;
    ForceFilter& force_filter = filters.force;
    RotationTracker& rotation = filters.rotation;
    Fingers& fingers = filters.fingers;
    IdleGate& idle = filters.idle;
    Ring interval_us, gpu_us;
    ~FluidRenderer() {
        if (sim) fluid_sim_destroy(sim);
        if (bench) fluid::stage0_destroy(bench);
    }
};

extern "C" {

const uint32_t FLUID_TOUCH_SLOTS = MAX_TOUCHES;

FluidRenderer* fluid_renderer_create(void* metal_layer, uint32_t width, uint32_t height, uint32_t particle_count,
                                     float sprite_radius, uint32_t bench_sweeps, float bench_spacing,
                                     uint32_t sim_substeps, uint32_t tracers, float particle_scale) {
    (void)metal_layer;
    (void)particle_count;
    (void)sprite_radius;
    if (bench_sweeps > 0) {
        // The stage-0 microbench (render.rs:2455-2466, 2861-2862): its own
        // spacing, 0.8 of the sim's unless given, clamped as the sim's is.
        if (fluid_sim_device_count() < 1) {
            fprintf(stderr, "fluid_renderer_create: no HIP device\n");
            return nullptr;
        }
        auto r = std::make_unique<FluidRenderer>();
        r->width = width;
        r->height = height;
        r->extent[0] = (float)width * 0.5f * METRES_PER_PIXEL;
        r->extent[1] = (float)height * 0.5f * METRES_PER_PIXEL;
        r->spacing = std::min(std::max(bench_spacing > 0.0f ? bench_spacing : 0.8f * SIM_SPACING, 0.4f * SIM_SPACING),
                              4.0f * SIM_SPACING);
        r->bench = fluid::stage0_create(r->extent, r->spacing, bench_sweeps);
        if (!r->bench) {
            fprintf(stderr, "fluid_renderer_create: %s\n", fluid_sim_last_error());
            return nullptr;
        }
        r->count = fluid::stage0_count(r->bench);
        return r.release();
    }
    if (sim_substeps == 0) {
        fprintf(stderr, "fluid_renderer_create: the sprite demo mode is not served by this library\n");
        return nullptr;
    }
    auto r = std::make_unique<FluidRenderer>();
    r->width = width;
    r->height = height;
    r->extent[0] = (float)width * 0.5f * METRES_PER_PIXEL;
    r->extent[1] = (float)height * 0.5f * METRES_PER_PIXEL;
    r->spacing = bench_spacing > 0.0f ? std::min(std::max(bench_spacing, 0.4f * SIM_SPACING), 4.0f * SIM_SPACING)
                                      : fluid_spacing_for(particle_scale, r->extent);
    r->max_substeps = sim_substeps;
    r->tracers = tracers;
    r->sim = build_sim(r->extent, r->spacing, tracers, &r->count);
    if (!r->sim) {
        fprintf(stderr, "fluid_renderer_create: %s\n", fluid_sim_last_error());
        return nullptr;
    }
    return r.release();
}

uint32_t fluid_renderer_frame(FluidRenderer* r, FluidVec3 gravity, FluidVec3 user_acceleration,
                              FluidVec3 rotation_rate, double now_ms) {
    if (!r) return 0;
    const float interval_us = (float)((now_ms - r->last_frame_ms) * 1000.0);
    const bool was_asleep = r->idle.sleeping();
    const float g[3] = {gravity.x, gravity.y, gravity.z};
    const float ua[3] = {user_acceleration.x, user_acceleration.y, user_acceleration.z};
    float body[3], force[3], omega[3], domega[3];
    fluid_body_force(g, ua, body);
    const float dev = r->force_filter.apply(body, force);
    const float rot[3] = {rotation_rate.x, rotation_rate.y, rotation_rate.z};
    r->rotation.apply(rot, interval_us / 1000000.0f, omega, domega);
    if (r->bench) {
        // The microbench encodes its sort and sweeps every frame and never
        // sleeps (render.rs:3042, 3164); the first frame validates.
        if (r->frames > 0 && interval_us < 500000.0f) r->interval_us.push(interval_us);
        r->last_frame_ms = now_ms;
        if (fluid::stage0_frame(r->bench) == FLUID_OK) r->gpu_us.push((float)(fluid::stage0_last_ms(r->bench) * 1e3));
        r->frames += 1;
        return 1;
    }
    if (r->idle.asleep(r->fingers.any_down(), force, dev, omega, r->stats[6])) {
        r->last_frame_ms = now_ms;
        r->idle_frames += 1;
        return 0;
    }
    if (r->frames > 0 && !was_asleep && interval_us < 500000.0f) r->interval_us.push(interval_us);
    r->last_frame_ms = now_ms;
    const float dt = r->frames == 0 ? 0.0f : std::min(std::max(interval_us / 1000000.0f, 0.0f), MAX_DT);
    const Touches touches = r->fingers.step(r->extent, dt);
    // The ramp chases the frame's own two ends from the last readback
    // (render.rs:3069-3077).
    r->ramp_live = r->ramp_look && r->ramp.follow(r->lens, r->stats, r->spacing, dt, r->ramp_ends);
    // From the one-frame-stale v_max, as the reference's readback gives it.
    r->substeps_used = dt > 0.0f ? fluid_substeps_for(dt, r->stats[6], r->spacing, r->max_substeps) : 0u;
    r->frame_seed += 1;
    if (dt > 0.0f && r->substeps_used > 0) {
        const auto started = std::chrono::steady_clock::now();
        const uint32_t n = r->substeps_used;
        const float dt_sub = dt / (float)n;
        const float v_clamp = 0.4f * r->spacing / dt_sub;
        uint8_t step[144], tracer_step[144];
        fluid_pack_step(force, omega, domega, dt_sub, v_clamp, 0, &touches.each[0][0], touches.count,
                        touches.radius, step);
        const float zero[3] = {0, 0, 0};
        fluid_pack_step(force, zero, zero, dt, 0.0f, r->frame_seed, nullptr, 0, 0.0f, tracer_step);
        // The glass look alone draws the tracers (render.rs:3041-3043).
        const bool strands = r->glass && r->tracers > 0;
        const uint32_t refine = fluid_refine_passes(dt_sub);
        std::memcpy(r->last_step, step, sizeof step);
        std::memcpy(r->last_tracer_step, strands ? tracer_step : step, sizeof step);
        r->last_refine = refine;
        fluid_sim_frame(r->sim, step, n, refine, strands ? tracer_step : nullptr);
        fluid_sim_read_stats(r->sim, r->stats);
        const float us = std::chrono::duration<float, std::micro>(std::chrono::steady_clock::now() - started).count();
        r->gpu_us.push(us);
    }
    r->frames += 1;
    return 1;
}

void fluid_renderer_set_particles(FluidRenderer* r, float scale) {
    if (!r || !r->sim) return;  // render.rs:2943-2946: the sim alone reseeds
    const float spacing = fluid_spacing_for(scale, r->extent);
    if (spacing == r->spacing) return;
    uint32_t count = 0;
    FluidSim* sim = build_sim(r->extent, spacing, r->tracers, &count);
    if (!sim) {
        fprintf(stderr, "fluid_renderer_set_particles: %s\n", fluid_sim_last_error());
        return;
    }
    fluid_sim_destroy(r->sim);
    r->sim = sim;
    r->spacing = spacing;
    r->count = count;
    std::memset(r->stats, 0, sizeof r->stats);
    r->idle = IdleGate();
}

uint32_t fluid_renderer_particles_at(const FluidRenderer* r, float scale) {
    if (!r || !r->sim) return 0;  // zero outside the sim (render.rs:2969)
    const float half[3] = {r->extent[0], r->extent[1], fluid_slab_depth() * 0.5f};
    uint32_t d[3];
    fluid_lattice_dims(fluid_spacing_for(scale, r->extent), half, 0.5f, d);
    return d[0] * d[1] * d[2];
}

void fluid_renderer_touch(FluidRenderer* r, uint32_t slot, float x, float y, bool down) {
    if (!r || slot >= (uint32_t)MAX_TOUCHES) return;
    r->fingers.down[slot] = down;
    r->fingers.at[slot][0] = x;
    r->fingers.at[slot][1] = y;
}

void fluid_renderer_set_look(FluidRenderer* r, bool flat, bool particles, FluidVec3 low, bool gradient,
                             FluidVec3 high, uint32_t lens) {
    (void)low;
    (void)high;
    if (!r) return;
    r->glass = !flat;
    (void)particles;  // Flat or Particles: both paint, neither is drawn here
    r->ramp_look = flat && gradient;
    r->lens = lens <= 4u ? lens : 0u;  // lib.rs:215-221: unknown codes paint velocity
    r->idle = IdleGate();
}

void fluid_renderer_resize(FluidRenderer* r, uint32_t width, uint32_t height) {
    if (!r || width == 0 || height == 0) return;
    r->width = width;
    r->height = height;
}

FluidRenderStats fluid_renderer_stats(const FluidRenderer* r) {
    FluidRenderStats s;
    std::memset(&s, 0, sizeof s);
    if (!r) return s;
    const float* f = r->stats;
    s.frames = r->frames;
    s.interval_p50_us = r->interval_us.percentile(0.5f);
    s.interval_p99_us = r->interval_us.percentile(0.99f);
    s.interval_max_us = r->interval_us.max();
    s.gpu_p50_us = r->gpu_us.percentile(0.5f);
    s.gpu_p99_us = r->gpu_us.percentile(0.99f);
    s.compression_avg = f[0];
    s.compression_max = f[1];
    s.density_min = f[2];
    s.density_max = f[3];
    s.pressure_min = f[4];
    s.pressure_max = f[5];
    s.v_max = f[6];
    s.temperature_min = f[7];
    s.temperature_max = f[8];
    s.clamp_count = (uint32_t)f[9];
    s.neighbour_overflow = (uint32_t)f[10];
    s.substeps = r->substeps_used;
    s.idle_frames = r->idle_frames;
    return s;
}

void fluid_renderer_destroy(FluidRenderer* r) { delete r; }

// ---- additions (include/fluid_sim.h) -------------------------------------------

FluidSim* fluid_renderer_sim(FluidRenderer* r) { return r ? r->sim : nullptr; }

int fluid_renderer_bench(FluidRenderer* r, int* scan_ok, int* density_ok, uint32_t* count, uint32_t* cells,
                         float* density, uint32_t* starts, float* pos4) {
    if (!r || !r->bench) return FLUID_ERR_ARG;
    fluid::stage0_validation(r->bench, scan_ok, density_ok);
    if (count) *count = fluid::stage0_count(r->bench);
    if (cells) *cells = fluid::stage0_cells(r->bench);
    return fluid::stage0_download(r->bench, density, starts, pos4);
}

int fluid_renderer_last_step(const FluidRenderer* r, uint8_t* step144, uint32_t* substeps, uint32_t* refine,
                             uint8_t* tracer_step144) {
    if (!r) return FLUID_ERR_ARG;
    if (step144) std::memcpy(step144, r->last_step, 144);
    if (tracer_step144) std::memcpy(tracer_step144, r->last_tracer_step, 144);
    if (substeps) *substeps = r->substeps_used;
    if (refine) *refine = r->last_refine;
    return FLUID_OK;
}

int fluid_renderer_ramp(const FluidRenderer* r, float* ends2) {
    if (!r || !r->ramp_live) return 0;
    if (ends2) {
        ends2[0] = r->ramp_ends[0];
        ends2[1] = r->ramp_ends[1];
    }
    return 1;
}

int fluid_lens_ends(uint32_t lens, const float* stats14, float spacing, float* ends2, float* floor) {
    if (!stats14) return FLUID_ERR_ARG;
    const uint32_t l = lens <= 4u ? lens : 0u;
    if (ends2) lens_ends(l, stats14, ends2);
    if (floor) *floor = lens_floor(l, spacing);
    return FLUID_OK;
}

FluidFilters* fluid_filters_create(void) { return new FluidFilters(); }
void fluid_filters_destroy(FluidFilters* f) { delete f; }

float fluid_filters_force(FluidFilters* f, const float body_force[3], float smooth[3]) {
    float tmp[3];
    return f && body_force ? f->force.apply(body_force, smooth ? smooth : tmp) : 0.0f;
}

void fluid_filters_rotation(FluidFilters* f, const float rate[3], float dt, float omega[3], float domega[3]) {
    if (!f || !rate) return;
    float o[3], d[3];
    f->rotation.apply(rate, dt, omega ? omega : o, domega ? domega : d);
}

void fluid_filters_touch(FluidFilters* f, uint32_t slot, float x, float y, bool down) {
    if (!f || slot >= (uint32_t)MAX_TOUCHES) return;
    f->fingers.down[slot] = down;
    f->fingers.at[slot][0] = x;
    f->fingers.at[slot][1] = y;
}

uint32_t fluid_filters_fingers(FluidFilters* f, const float extent[2], float dt, float* touches20, float* radius) {
    if (!f || !extent) return 0;
    const Touches t = f->fingers.step(extent, dt);
    if (touches20) std::memcpy(touches20, &t.each[0][0], sizeof t.each);
    if (radius) *radius = t.radius;
    return t.count;
}

bool fluid_filters_asleep(FluidFilters* f, bool touched, const float smooth[3], float dev, const float omega[3],
                          float v_max) {
    if (!f || !smooth || !omega) return false;
    return f->idle.asleep(touched, smooth, dev, omega, v_max);
}

uint32_t fluid_filters_still_frames(void) { return IdleGate::STILL_FRAMES; }

}  // extern "C"
