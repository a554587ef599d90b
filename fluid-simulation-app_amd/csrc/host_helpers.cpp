// The reference's CPU substrate of the step, in C++: grid arithmetic,
// kernel, seeding, the byte packers (crates/fluid-core/src/sim.rs:1-350),
// the body force (lib.rs:34-36) and the substep schedule (render.rs:57-106,
// 1362-1380). Pure host code; exported through include/fluid_sim.h.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/fluid_sim.h"

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float WORLD_SCALE = 4.0f;                    // lib.rs:13
constexpr float STANDARD_GRAVITY = 9.80665f;           // lib.rs:5
constexpr float SLAB_DEPTH = WORLD_SCALE * 0.00765f;   // sim.rs:10
constexpr float SIM_SPACING = WORLD_SCALE * 0.0025f;   // render.rs:1353
constexpr float SUBSTEP_PER_SPACING = 0.42f;           // render.rs:55
constexpr float REFINE_SHORT_DT = 0.00105f;            // render.rs:68
constexpr float NOMINAL_FRAME = 1.0f / 120.0f;         // render.rs:81
constexpr float CHEAP_RUNG_COST = 0.65f;               // render.rs:87

// Rust's saturating float -> u32 `as` conversion.
uint32_t as_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}

void put_f32(uint8_t* raw, size_t off, float v) { std::memcpy(raw + off, &v, 4); }
void put_u32(uint8_t* raw, size_t off, uint32_t v) { std::memcpy(raw + off, &v, 4); }

}  // namespace

extern "C" {

uint32_t fluid_hash(uint32_t x) {  // sim.rs:318-323
    uint32_t h = x * 0x9E3779B9u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    return h ^ (h >> 13);
}

void fluid_grid_new(const float half[3], float support, FluidGrid* out) {  // sim.rs:24-32
    for (int i = 0; i < 3; ++i) {
        out->dims[i] = as_u32(std::ceil(2.0f * half[i] / support)) + 2u;
        out->min[i] = -half[i] - support;
    }
    out->cell = support;
}

uint32_t fluid_grid_cell_of(const FluidGrid* g, const float pos[3]) {  // sim.rs:38-43
    uint32_t c[3];
    for (int i = 0; i < 3; ++i) c[i] = std::min(as_u32((pos[i] - g->min[i]) / g->cell), g->dims[i] - 1u);
    return (c[2] * g->dims[1] + c[1]) * g->dims[0] + c[0];
}

float fluid_kernel(float r, float h) {  // sim.rs:49-60
    const float q = r / h;
    const float sigma = 1.0f / (kPi * h * h * h);
    if (q < 1.0f) return sigma * (1.0f - 1.5f * q * q * (1.0f - 0.5f * q));
    if (q < 2.0f) {
        const float t = 2.0f - q;
        return sigma * 0.25f * t * t * t;
    }
    return 0.0f;
}

void fluid_pack_sim_params(const FluidGrid* g, uint32_t count, float h, float mass, float rho0, uint8_t* out64) {
    // sim.rs:330-350; rho0 is a parameter here (the reference packs REST_DENSITY).
    std::memset(out64, 0, 64);
    for (int i = 0; i < 3; ++i) put_f32(out64, (size_t)i * 4, g->min[i]);
    put_f32(out64, 12, g->cell);
    for (int i = 0; i < 3; ++i) put_u32(out64, 16 + (size_t)i * 4, g->dims[i]);
    put_u32(out64, 28, count);
    put_f32(out64, 32, h);
    put_f32(out64, 36, mass);
    put_f32(out64, 40, rho0);
    const float sigma = 1.0f / (kPi * h * h * h);
    put_f32(out64, 44, 1.0f / h);
    put_f32(out64, 48, sigma);
    put_f32(out64, 52, sigma / h);
    put_f32(out64, 56, 4.0f * h * h);
}

void fluid_pack_step(const float force[3], const float omega[3], const float domega[3], float dt, float v_clamp,
                     uint32_t seed, const float* touches20, uint32_t touch_n, float touch_r, uint8_t* out144) {
    // sim.rs:215-252
    std::memset(out144, 0, 144);
    const float head[11] = {force[0], force[1], force[2], dt, omega[0], omega[1], omega[2], v_clamp,
                            domega[0], domega[1], domega[2]};
    for (int i = 0; i < 11; ++i) put_f32(out144, (size_t)i * 4, head[i]);
    put_u32(out144, 44, seed);
    put_f32(out144, 48, touch_r);
    put_u32(out144, 52, touch_n);
    if (touches20)
        for (int i = 0; i < 20; ++i) put_f32(out144, 64 + (size_t)i * 4, touches20[i]);
}

void fluid_body_force(const float gravity[3], const float user_acceleration[3], float out[3]) {  // lib.rs:34-36
    for (int i = 0; i < 3; ++i) out[i] = STANDARD_GRAVITY * (gravity[i] + user_acceleration[i]);
}

void fluid_lattice_dims(float spacing, const float half[3], float fill, uint32_t out[3]) {  // sim.rs:65-72
    const float size[3] = {2.0f * half[0] - spacing, (2.0f * half[1] - spacing) * fill, 2.0f * half[2] - spacing};
    for (int i = 0; i < 3; ++i) out[i] = std::max(as_u32(size[i] / spacing), 1u);
}

uint32_t fluid_seed_slab(float spacing, const float half[3], float fill, float* out4) {  // sim.rs:77-100
    uint32_t n[3];
    fluid_lattice_dims(spacing, half, fill, n);
    const uint32_t total = n[0] * n[1] * n[2];
    if (!out4) return total;
    const float origin[3] = {-0.5f * (float)n[0] * spacing, -half[1] + 0.5f * spacing,
                             -0.5f * (float)n[2] * spacing};
    auto jitter = [](uint32_t h) { return ((float)h / 4294967296.0f - 0.5f) * 0.2f; };
    for (uint32_t iz = 0; iz < n[2]; ++iz)
        for (uint32_t iy = 0; iy < n[1]; ++iy)
            for (uint32_t ix = 0; ix < n[0]; ++ix) {
                const uint32_t i = (iz * n[1] + iy) * n[0] + ix;
                float* o = out4 + (size_t)i * 4;
                o[0] = origin[0] + ((float)ix + 0.5f + jitter(fluid_hash(i * 3))) * spacing;
                o[1] = origin[1] + ((float)iy + jitter(fluid_hash(i * 3 + 1))) * spacing;
                o[2] = origin[2] + ((float)iz + 0.5f + jitter(fluid_hash(i * 3 + 2))) * spacing;
                o[3] = 0.0f;
            }
    return total;
}

uint16_t fluid_half_bits(float x) {  // sim.rs:139-151 (finite, below 65504)
    uint32_t xb;
    std::memcpy(&xb, &x, 4);
    const uint16_t sign = (uint16_t)((xb >> 16) & 0x8000u);
    const float a = std::fabs(x);
    uint32_t ab;
    std::memcpy(&ab, &a, 4);
    const uint32_t small_bits = 113u << 23;
    float small;
    std::memcpy(&small, &small_bits, 4);
    if (a < small) {
        const uint32_t magic_bits = 126u << 23;
        float magic;
        std::memcpy(&magic, &magic_bits, 4);
        const float s = a + magic;
        uint32_t sb;
        std::memcpy(&sb, &s, 4);
        return (uint16_t)(sign | (uint16_t)(sb - magic_bits));
    }
    return (uint16_t)(sign | (uint16_t)(((ab + 0x0fffu + ((ab >> 13) & 1u)) - (112u << 23)) >> 13));
}

void fluid_pack_tracer(const float pos[3], float speed, const float extent[2], uint32_t out[2]) {  // sim.rs:127-133
    auto unorm = [](float v, float e) {
        return as_u32(std::min(std::max(v / e * 0.5f + 0.5f, 0.0f), 1.0f) * 65535.0f + 0.5f);
    };
    out[0] = unorm(pos[0], extent[0]) | (unorm(pos[1], extent[1]) << 16);
    out[1] = (uint32_t)fluid_half_bits(pos[2]) | ((uint32_t)fluid_half_bits(speed) << 16);
}

void fluid_seed_tracers(uint32_t count, const float half3[3], float fill, uint32_t* out2) {  // sim.rs:105-122
    const float half[3] = {half3[0] - 0.001f, half3[1] - 0.001f, half3[2] - 0.001f};
    const float extent[2] = {half3[0], half3[1]};
    auto unit = [](uint32_t h) { return (float)h / 4294967296.0f; };
    for (uint32_t i = 0; i < count; ++i) {
        const float pos[3] = {(unit(fluid_hash(i * 3)) * 2.0f - 1.0f) * half[0],
                              -half[1] + unit(fluid_hash(i * 3 + 1)) * 2.0f * half[1] * fill,
                              (unit(fluid_hash(i * 3 + 2)) * 2.0f - 1.0f) * half[2]};
        fluid_pack_tracer(pos, 0.0f, extent, out2 + (size_t)i * 2);
    }
}

uint32_t fluid_refine_passes(float dt_sub) { return dt_sub > REFINE_SHORT_DT ? 5u : 2u; }  // render.rs:77-79

uint32_t fluid_substeps_for(float dt, float v_max, float spacing, uint32_t cap) {  // render.rs:62-106
    const uint32_t floor_n = std::min(as_u32(std::ceil(dt / (SUBSTEP_PER_SPACING * spacing))), 8u);
    uint32_t n = std::max(as_u32(std::ceil(dt * v_max / (0.4f * spacing))), floor_n);
    const uint32_t cheap = as_u32(std::ceil(NOMINAL_FRAME / REFINE_SHORT_DT));
    const bool jump = fluid_refine_passes(dt / (float)cheap) == 2u && n < cheap &&
                      (float)n > CHEAP_RUNG_COST * (float)cheap;
    if (jump) n = cheap;
    return std::min(n, cap);
}

float fluid_spacing_for(float scale, const float extent[2]) {  // render.rs:1362-1380
    const float half[3] = {extent[0], extent[1], SLAB_DEPTH * 0.5f};
    auto count = [&](float spacing) {
        uint32_t d[3];
        fluid_lattice_dims(spacing, half, 0.5f, d);
        return d[0] * d[1] * d[2];
    };
    const float target = scale * (float)count(SIM_SPACING);
    const float ideal = SIM_SPACING / std::cbrt(scale);
    float best_miss = INFINITY, best_off = INFINITY, best = SIM_SPACING;
    for (int permille = 400; permille <= 4000; permille += 5) {
        const float spacing = SIM_SPACING * (float)permille / 1000.0f;
        const float miss = std::fabs(std::log((float)count(spacing) / target));
        const float off = std::fabs(spacing - ideal);
        if (miss < best_miss || (miss == best_miss && off < best_off)) {
            best_miss = miss;
            best_off = off;
            best = spacing;
        }
    }
    return best;
}

float fluid_slab_depth(void) { return SLAB_DEPTH; }

}  // extern "C"
