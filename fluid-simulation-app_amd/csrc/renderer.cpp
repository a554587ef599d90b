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

struct FluidRenderer {
    FluidSim* sim = nullptr;
    float extent[2] = {0, 0};
    float spacing = 0.0f;
    uint32_t max_substeps = 0, tracers = 0, count = 0, width = 0, height = 0;
    uint32_t substeps_used = 0, frame_seed = 0;
    uint64_t frames = 0, idle_frames = 0;
    double last_frame_ms = 0.0;
    float stats[FLUID_SIM_STATS] = {};
    bool glass = true;  // the reference's default look, Look::Glass
    ForceFilter force_filter;
    RotationTracker rotation;
    Fingers fingers;
    IdleGate idle;
    Ring interval_us, gpu_us;
    ~FluidRenderer() {
        if (sim) fluid_sim_destroy(sim);
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
        fprintf(stderr, "fluid_renderer_create: the stage-0 microbench mode is not served by this library\n");
        return nullptr;
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
    if (r->idle.asleep(r->fingers.any_down(), force, dev, omega, r->stats[6])) {
        r->last_frame_ms = now_ms;
        r->idle_frames += 1;
        return 0;
    }
    if (r->frames > 0 && !was_asleep && interval_us < 500000.0f) r->interval_us.push(interval_us);
    r->last_frame_ms = now_ms;
    const float dt = r->frames == 0 ? 0.0f : std::min(std::max(interval_us / 1000000.0f, 0.0f), MAX_DT);
    const Touches touches = r->fingers.step(r->extent, dt);
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
        fluid_sim_frame(r->sim, step, n, fluid_refine_passes(dt_sub), strands ? tracer_step : nullptr);
        fluid_sim_read_stats(r->sim, r->stats);
        const float us = std::chrono::duration<float, std::micro>(std::chrono::steady_clock::now() - started).count();
        r->gpu_us.push(us);
    }
    r->frames += 1;
    return 1;
}

void fluid_renderer_set_particles(FluidRenderer* r, float scale) {
    if (!r) return;
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
    if (!r) return 0;
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
    (void)particles;
    (void)low;
    (void)gradient;
    (void)high;
    (void)lens;
    if (!r) return;
    r->glass = !flat;
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

}  // extern "C"
