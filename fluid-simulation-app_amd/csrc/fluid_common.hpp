// Shared definitions of the MI355X DFSPH step: the parameter blocks the
// kernels take by value, and the device math of the step.
//
// Byte layouts follow the reference's SimParams (crates/fluid-core/src/sim.rs:325-350,
// sim_solve.wgsl:11-23) and Step (sim.rs:200-252, sim_solve.wgsl:29-41).
// Every expression below keeps the reference shader's operand order, and the
// library is built with -ffp-contract=off and IEEE division and square root,
// so each float result is that of single precision evaluated in that order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fluid {

// ---- byte blocks of the reference's parameter API -----------------------

struct SimParamsBlock {  // 64 B, sim.rs:330-350
    float box_min[3];
    float cell;
    uint32_t dims[3];
    uint32_t count;
    float h, mass, rho0, inv_h, sigma, sigma_h, support_sq, pad;
};
static_assert(sizeof(SimParamsBlock) == 64, "SimParams is 64 bytes");

struct StepBlock {  // 144 B, sim.rs:215-252
    float force[3];
    float dt;
    float omega[3];
    float v_clamp;
    float domega[3];
    uint32_t seed;
    float touch_r;
    uint32_t touch_n;
    uint32_t pad[2];
    float touches[5][4];
};
static_assert(sizeof(StepBlock) == 144, "Step is 144 bytes");

// ---- constants of the method, sim_solve.wgsl:84-159, sim_tracers.wgsl:39-52

constexpr float XSPH_RATE = 6.0f;
constexpr float TOUCH_RATE = 40.0f;
constexpr float LENS_RATE = 20.0f;
constexpr float K_NEAR = 3000.0f;
constexpr float DYNAMIC_VISCOSITY = 1.002e-3f;
constexpr float HEAT_CAPACITY = 4184.0f;
constexpr float CONDUCTIVITY = 0.598f;
constexpr float EXPANSION = 2.07e-4f;
constexpr float SIGMA_WATER = 0.0728f;
constexpr float LATTICE_SIGMA = 0.8665f;
constexpr float COS_CONTACT = -0.342f;
constexpr float FIXED = 65536.0f;
constexpr float TAU = 3.0f;
constexpr float TAU_STRAY = 0.25f;
constexpr float T_CHARGE = 4.0f;
constexpr float HUGE_F = 1e30f;

// ---- what the kernels take by value --------------------------------------

// Physical coefficients the reference fixes at compile time
// (sim_solve.wgsl:135-137, 161-172, 657), as runtime values: the dynamic
// viscosity (a parameter of the adjoint), and switches for the terms the
// adjoint's differentiable profile leaves out (near pressure, cohesion,
// adhesion, XSPH). The defaults are the reference's.
struct Physics {
    float mu = DYNAMIC_VISCOSITY;
    float k_near = K_NEAR;
    float tension_scale = 1.0f;   // multiplies tension_gamma
    float adhesion_scale = 1.0f;  // multiplies adhesion_beta
    float xsph_rate = XSPH_RATE;
};

// Scene constants: the SimParams block plus the pair-loop constants the
// shader hoists (sim_solve.wgsl:161-172, 534-537), evaluated once on the host
// in single precision in the shader's operand order.
struct SceneK {
    float box_min[3];
    float cell;
    uint32_t dims[3];
    uint32_t count;
    float h, mass, rho0, inv_h, sigma, sigma_h, support_sq;
    float wall_lo[3], wall_hi[3];
    float inv_h_wall;  // 1.0 / h as wall_grad_sum and wall_adh_sum evaluate it
    float coh_c, coh_k, coh_near, gamma, beta;
    float visc2, heat_cond, heat_visc, eps_h2;
    float k_near;
    // Where the scene lives in a batch (0, 0, count, C for a lone scene): its
    // cells are [cell_offset, cell_offset + cells) of the shared histogram,
    // plus one padding cell after them that the slots past its particles
    // sort into; its particles are slots [slot_begin, slot_end).
    uint32_t cell_offset, cells, slot_begin, slot_end;
};

// Substep constants: the Step block plus the exponentials the shader
// evaluates per invocation, which are uniform over a substep.
struct StepK {
    float force[3];
    float dt;
    float omega[3];
    float v_clamp;
    float domega[3];
    uint32_t seed;
    float touch_r;
    uint32_t touch_n;
    float touches[5][4];
    float c_xsph;    // 1 - exp(-XSPH_RATE dt), sim_solve.wgsl:657
    float chase;     // 1 - exp(-LENS_RATE dt), sim_solve.wgsl:741
    float decay;     // exp(-dt / T_CHARGE), sim_tracers.wgsl:182
    float dt_guard;  // max(dt, 1e-9), sim_solve.wgsl:516
    uint32_t refine;    // density-solve passes of this substep (render.rs:77-79)
    uint32_t active;    // 0: the scene sits this substep out (batched steps)
};

SceneK make_scene(const SimParamsBlock& p, const Physics& phys = Physics());
StepK make_step(const StepBlock& s, const Physics& phys = Physics());

// ---- device math (the reference's WGSL helpers) --------------------------

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 scl(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float wmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float wmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float wmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// float -> u32 as WGSL and Rust convert: truncation, negatives and NaN to 0,
// saturation at the top.
__device__ __forceinline__ uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
__device__ __forceinline__ int32_t sat_i32(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)0x80000000u;
    return (int32_t)f;
}

// sim_grid.wgsl:36-42: the clamped, truncated cell coordinate.
__device__ __forceinline__ uint32_t cell_coord(float p, float lo, float cell, uint32_t dim) {
    uint32_t c = sat_u32((p - lo) / cell);
    return c < dim - 1u ? c : dim - 1u;
}

// e^x for the fingers' per-particle blend (sim_solve.wgsl:124), evaluated
// with a fixed sequence of IEEE operations so that host code following the
// same sequence gets the same bits: Cody-Waite reduction x = k ln2 + r with
// ln2 split into a 16-bit head (k * head is exact) and a tail, a degree-7
// Taylor polynomial for e^r on |r| <= ln2/2 (truncation below 2^-27), and an
// exact scaling by 2^k. Valid for x in [-87, 88]; WGSL allows exp() 3 + 2|x|
// ULP, which this meets.
__device__ __forceinline__ float exp_fixed(float x) {
    x = wmin(wmax(x, -87.0f), 88.0f);
    const float k = rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-06f;
    const float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666672f + r * (0.0416666679f + r * (0.00833333377f +
                    r * (0.00138888892f + r * 0.000198412701f))))));
    return ldexpf(p, (int)k);
}

// sim_solve.wgsl:176-182
__device__ __forceinline__ float kernel_q(const SceneK& k, float q) {
    if (q < 1.0f) return k.sigma * (1.0f - 1.5f * q * q * (1.0f - 0.5f * q));
    float t = 2.0f - q;
    return k.sigma * 0.25f * t * t * t;
}
// Correctly rounded square root and quotient for the pair loops, in fewer
// instructions than the general ones. The general sequences are the
// hardware estimate and a residual correction, wrapped in operand scaling
// and special-value fixups; here these are identity, so the cores alone give
// the same bits:
// - sqrt_rn(x) for x = 0 or x >= 2^-96 (a positive x below that, which only
//   coincident particles give, returns some value under 2^-47, which every
//   caller treats as it treats 0);
// - div_rn(a, b) for a normal b and |a|, |b|, |a / b| well inside 2^+-64
//   (or a = 0, which returns +0: the callers add such a quotient, times
//   finite factors, to sums that start at +0).
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __int_as_float(__float_as_int(s) - 1), up = __int_as_float(__float_as_int(s) + 1);
    float r = __builtin_fmaf(-dn, s, x) <= 0.0f ? dn : s;
    r = __builtin_fmaf(-up, s, x) > 0.0f ? up : r;
    return r;
}
__device__ __forceinline__ float div_rn(float a, float b) {
    const float r0 = __builtin_amdgcn_rcpf(b);
    const float r1 = __builtin_fmaf(__builtin_fmaf(-b, r0, 1.0f), r0, r0);
    const float q0 = a * r1;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-b, q0, a), r1, q0);
    return __builtin_fmaf(__builtin_fmaf(-b, q1, a), r1, q1);
}

// sim_solve.wgsl:184-196: dW/dq, the numerator of the scalar dW/r that
// multiplies x.
__device__ __forceinline__ float grad_dw(const SceneK& k, float q) {
    if (q < 1.0f) return k.sigma_h * (-3.0f * q + 2.25f * q * q);
    float t = 2.0f - q;
    return k.sigma_h * (-0.75f * t * t);
}
__device__ __forceinline__ float grad_scale(const SceneK& k, float r, float q) { return grad_dw(k, q) / r; }
__device__ __forceinline__ f3 grad_q(const SceneK& k, f3 x, float r, float q) {
    if (r < 1e-9f) return mk(0.0f, 0.0f, 0.0f);
    return scl(grad_scale(k, r, q), x);
}
// sim_solve.wgsl:201-210
__device__ __forceinline__ float cohesion(float r, float c, float kk, float nearc) {
    if (r >= c || r < 1e-9f) return 0.0f;
    float a = (c - r) * (c - r) * (c - r) * r * r * r;
    if (2.0f * r > c) return kk * a;
    return kk * (2.0f * a - nearc);
}
// A dual number (value, derivative) for the adjoint's derivatives of the
// wall polynomials: the templates below evaluate at float for the step and
// at Dn for the step's derivative, the same operations in the same order.
struct Dn {
    float v, d;
    __device__ __forceinline__ Dn(float x = 0.0f, float dx = 0.0f) : v(x), d(dx) {}
};
__device__ __forceinline__ Dn operator+(Dn a, Dn b) { return Dn(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dn operator-(Dn a, Dn b) { return Dn(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dn operator*(Dn a, Dn b) { return Dn(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dn operator/(Dn a, float b) { return Dn(a.v / b, a.d / b); }
__device__ __forceinline__ Dn operator+(float a, Dn b) { return Dn(a + b.v, b.d); }
__device__ __forceinline__ Dn operator-(float a, Dn b) { return Dn(a - b.v, -b.d); }
__device__ __forceinline__ Dn operator*(float a, Dn b) { return Dn(a * b.v, a * b.d); }
__device__ __forceinline__ Dn operator+(Dn a, float b) { return Dn(a.v + b, a.d); }
__device__ __forceinline__ Dn operator-(Dn a, float b) { return Dn(a.v - b, a.d); }
__device__ __forceinline__ Dn operator*(Dn a, float b) { return Dn(a.v * b, a.d * b); }
__device__ __forceinline__ Dn& operator+=(Dn& a, Dn b) { return a = a + b; }
__device__ __forceinline__ Dn& operator-=(Dn& a, Dn b) { return a = a - b; }
__device__ __forceinline__ bool operator>=(Dn a, float b) { return a.v >= b; }
__device__ __forceinline__ bool operator<(Dn a, float b) { return a.v < b; }
__device__ __forceinline__ float value(float x) { return x; }
__device__ __forceinline__ float value(Dn x) { return x.v; }

// sim_solve.wgsl:213-222
template <class N>
__device__ __forceinline__ N wall_density_t(N t) {
    if (t >= 2.0f) return N(0.0f);
    N t3 = t * t * t;
    if (t < 1.0f) return 0.5f - 0.7f * t + t3 * ((float)(1.0 / 3.0) + t * t * (-0.15f + 0.05f * t));
    return (float)(8.0 / 15.0) - 0.8f * t + t3 * ((float)(2.0 / 3.0) + t * (-0.5f + t * (0.15f - t / 60.0f)));
}
__device__ __forceinline__ float wall_density(float t) { return wall_density_t<float>(t); }
// sim_solve.wgsl:226-235
template <class N>
__device__ __forceinline__ N wall_gradient_t(N t) {
    if (t >= 2.0f) return N(0.0f);
    N t2 = t * t;
    if (t < 1.0f) return 0.7f + t2 * (-1.0f + t2 * (0.75f - t * 0.3f));
    return 0.8f + t2 * (-2.0f + t * (2.0f + t * (-0.75f + t * 0.1f)));
}
__device__ __forceinline__ float wall_gradient(float t) { return wall_gradient_t<float>(t); }
// sim_solve.wgsl:249-270 — the reference's fitted coefficient table.
template <class N>
__device__ __forceinline__ N wedge_t(N t1, N t2) {
    if (t1 * t1 + t2 * t2 >= 4.0f) return N(0.0f);
    N c0 = 2.5005807e-01f + t2 * (-3.4681153e-01f + t2 * (-4.0986882e-02f + t2 * (3.0301620e-01f + t2 * (-1.9202736e-01f + t2 * (4.6947582e-02f + t2 * (-3.8543515e-03f))))));
    N c1 = -3.4681153e-01f + t2 * (4.8587248e-01f + t2 * (1.0092180e-02f + t2 * (-2.9978363e-01f + t2 * (1.5320335e-01f + t2 * (-2.2544282e-02f)))));
    N c2 = -4.0986882e-02f + t2 * (1.0092180e-02f + t2 * (-5.4145610e-02f + t2 * (7.9597369e-02f + t2 * (-2.3350373e-02f))));
    N c3 = 3.0301620e-01f + t2 * (-2.9978363e-01f + t2 * (7.9597369e-02f + t2 * (-8.2431946e-03f)));
    N c4 = -1.9202736e-01f + t2 * (1.5320335e-01f + t2 * (-2.3350373e-02f));
    N c5 = 4.6947582e-02f + t2 * (-2.2544282e-02f);
    return c0 + t1 * (c1 + t1 * (c2 + t1 * (c3 + t1 * (c4 + t1 * (c5 + t1 * (-3.8543515e-03f))))));
}
__device__ __forceinline__ float wedge(float t1, float t2) { return wedge_t<float>(t1, t2); }
// sim_solve.wgsl:272-293
template <class N>
__device__ __forceinline__ N wedge_d_t(N t1, N t2) {
    if (t1 * t1 + t2 * t2 >= 4.0f) return N(0.0f);
    N c0 = -3.4611994e-01f + t2 * (4.4757673e-01f + t2 * (1.1932939e-01f + t2 * (-4.6460101e-01f + t2 * (2.7637253e-01f + t2 * (-6.2846025e-02f + t2 * (4.3871652e-03f))))));
    N c1 = -6.3622823e-02f + t2 * (2.0521006e-01f + t2 * (-3.7311339e-01f + t2 * (3.7508155e-01f + t2 * (-1.6629713e-01f + t2 * (2.5030129e-02f)))));
    N c2 = 7.8352497e-01f + t2 * (-1.2658344e+00f + t2 * (5.8898945e-01f + t2 * (-1.0772903e-01f + t2 * (1.3535777e-02f))));
    N c3 = -4.9470084e-01f + t2 * (9.1302071e-01f + t2 * (-3.4330355e-01f + t2 * (2.4034609e-02f)));
    N c4 = -2.5605481e-02f + t2 * (-2.0067981e-01f + t2 * (5.9854468e-02f));
    N c5 = 8.8898623e-02f + t2 * (3.9914838e-03f);
    return c0 + t1 * (c1 + t1 * (c2 + t1 * (c3 + t1 * (c4 + t1 * (c5 + t1 * (-1.7788321e-02f))))));
}
__device__ __forceinline__ float wedge_d(float t1, float t2) { return wedge_d_t<float>(t1, t2); }
// sim_solve.wgsl:334-340
__device__ __forceinline__ float wall_adhesion(float t) {
    float u = wmin(wmax(2.0f - t, 0.0f), 2.0f);
    return u * u * (1.7847015e-3f + u * (3.3392196e-3f + u * (-4.3394677e-3f + u * (1.6711868e-3f + u * (-2.1788750e-4f)))));
}
template <class N>
__device__ __forceinline__ N wedge_dsum_t(N t, N a, N b, N c, N d) {
    return wedge_d_t(t, a) + wedge_d_t(t, b) + wedge_d_t(t, c) + wedge_d_t(t, d);
}
__device__ __forceinline__ float wedge_dsum(float t, float a, float b, float c, float d) {
    return wedge_dsum_t<float>(t, a, b, c, d);
}
// sim_solve.wgsl:299-304, components (lx, ly, lz) = tl, (hx, hy, hz) = th
template <class N>
__device__ __forceinline__ N wedge_sum_t(N lx, N ly, N lz, N hx, N hy, N hz) {
    N w = wedge_t(lx, ly) + wedge_t(lx, hy) + wedge_t(hx, ly) + wedge_t(hx, hy);
    w += wedge_t(lx, lz) + wedge_t(lx, hz) + wedge_t(hx, lz) + wedge_t(hx, hz);
    w += wedge_t(ly, lz) + wedge_t(ly, hz) + wedge_t(hy, lz) + wedge_t(hy, hz);
    return w;
}
__device__ __forceinline__ float wedge_sum(f3 tl, f3 th) {
    return wedge_sum_t<float>(tl.x, tl.y, tl.z, th.x, th.y, th.z);
}
// sim_solve.wgsl:310-328
template <class N>
__device__ __forceinline__ void wall_grad_sum_t(const SceneK& k, N px, N py, N pz, N& gx, N& gy, N& gz) {
    const float ih = k.inv_h_wall;
    const N lx = (px - k.wall_lo[0]) * ih, ly = (py - k.wall_lo[1]) * ih, lz = (pz - k.wall_lo[2]) * ih;
    const N hx = (k.wall_hi[0] - px) * ih, hy = (k.wall_hi[1] - py) * ih, hz = (k.wall_hi[2] - pz) * ih;
    const float scale = k.rho0 * ih;
    gx = N(0.0f);
    gy = N(0.0f);
    gz = N(0.0f);
    gx -= scale * (wall_gradient_t(lx) + wedge_dsum_t(lx, ly, hy, lz, hz));
    gx += scale * (wall_gradient_t(hx) + wedge_dsum_t(hx, ly, hy, lz, hz));
    gy -= scale * (wall_gradient_t(ly) + wedge_dsum_t(ly, lx, hx, lz, hz));
    gy += scale * (wall_gradient_t(hy) + wedge_dsum_t(hy, lx, hx, lz, hz));
    gz -= scale * (wall_gradient_t(lz) + wedge_dsum_t(lz, lx, hx, ly, hy));
    gz += scale * (wall_gradient_t(hz) + wedge_dsum_t(hz, lx, hx, ly, hy));
}
__device__ __forceinline__ f3 wall_grad_sum(const SceneK& k, f3 pos) {
    f3 g;
    wall_grad_sum_t<float>(k, pos.x, pos.y, pos.z, g.x, g.y, g.z);
    return g;
}
// The wall fill of density_div (sim_solve.wgsl:485-493): the kernel mass
// behind the six walls, less the doubly counted wedges.
template <class N>
__device__ __forceinline__ N wall_fill_t(const SceneK& k, N px, N py, N pz) {
    const N lx = (px - k.wall_lo[0]) * k.inv_h, ly = (py - k.wall_lo[1]) * k.inv_h, lz = (pz - k.wall_lo[2]) * k.inv_h;
    const N hx = (k.wall_hi[0] - px) * k.inv_h, hy = (k.wall_hi[1] - py) * k.inv_h, hz = (k.wall_hi[2] - pz) * k.inv_h;
    N fill = N(0.0f);
    fill += wall_density_t(lx) + wall_density_t(hx);
    fill += wall_density_t(ly) + wall_density_t(hy);
    fill += wall_density_t(lz) + wall_density_t(hz);
    fill -= wedge_sum_t(lx, ly, lz, hx, hy, hz);
    return fill;
}
// Whether a wall term can be nonzero at pos. Every wall term (density,
// wedges, gradient, adhesion) of a face is 0 once its t = distance x 1/h is
// at least 2; this is false only when the nearest face's distance times the
// smaller of the two 1/h scalings the terms use rounds to at least
// 2 (1 + 2^-22), so every t, under either scaling, rounds to at least 2 and
// skipping the terms gives the same +0 they would.
__device__ __forceinline__ bool near_walls(const SceneK& k, f3 p) {
    const float d = wmin(wmin(wmin(p.x - k.wall_lo[0], k.wall_hi[0] - p.x), wmin(p.y - k.wall_lo[1], k.wall_hi[1] - p.y)),
                         wmin(p.z - k.wall_lo[2], k.wall_hi[2] - p.z));
    return d * wmin(k.inv_h, k.inv_h_wall) < 2.000001f;
}
// sim_solve.wgsl:349-361
__device__ __forceinline__ f3 wall_adh_sum(const SceneK& k, f3 pos) {
    float ih = k.inv_h_wall;
    f3 a = mk(0.0f, 0.0f, 0.0f);
    a.x -= wall_adhesion((pos.x - k.wall_lo[0]) * ih);
    a.x += wall_adhesion((k.wall_hi[0] - pos.x) * ih);
    a.y -= wall_adhesion((pos.y - k.wall_lo[1]) * ih);
    a.y += wall_adhesion((k.wall_hi[1] - pos.y) * ih);
    a.z -= wall_adhesion((pos.z - k.wall_lo[2]) * ih);
    a.z += wall_adhesion((k.wall_hi[2] - pos.z) * ih);
    return a;
}

// lane_fold (sim_solve.wgsl:398-402) over the 4 lanes of a quad: xor 2, then
// xor 1 (every lane ends with (p0 + p2) + (p1 + p3)). DPP quad permutations,
// no LDS; every lane of the wave must take part.
template <int CTRL>
__device__ __forceinline__ float quad_perm(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float fold4(float v) {
    const float t = v + quad_perm<0x4E>(v);  // [2, 3, 0, 1]: lane ^ 2
    return t + quad_perm<0xB1>(t);           // [1, 0, 3, 2]: lane ^ 1
}
__device__ __forceinline__ f3 fold4(f3 v) { return mk(fold4(v.x), fold4(v.y), fold4(v.z)); }
// Lane K's value in every lane of its quad (quad_perm [K, K, K, K]).
template <int K>
__device__ __forceinline__ float quad_bcast(float v) {
    return quad_perm<K * 0x55>(v);
}
// A float4 spread over a quad (lane l holds component l), whole in every
// lane. The four lanes of a particle share the loads of its own fields this
// way: one register a lane where a whole float4 would take four.
__device__ __forceinline__ float4 quad_float4(float v) {
    return make_float4(quad_bcast<0>(v), quad_bcast<1>(v), quad_bcast<2>(v), quad_bcast<3>(v));
}
__device__ __forceinline__ float lane_part(const float4* a, uint32_t i, uint32_t lane) {
    return reinterpret_cast<const float*>(a)[(size_t)i * 4 + lane];
}
// base[i] through a 32-bit byte offset (i * sizeof(T) < 2^32, which the
// allocations guarantee): the load form that keeps a uniform base in scalar
// registers and takes the offset from one vector register, with no 64-bit
// address arithmetic a lane.
template <class T>
__device__ __forceinline__ T at32(const T* base, uint32_t i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ T* at32p(T* base, uint32_t i) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(T));
}

// Inclusive prefix sum over the 64 lanes of a wave: four row shifts build
// the 16-lane row scans, then two row broadcasts (lane 15 into rows 1 and 3,
// lane 31 into rows 2 and 3) carry the row totals across the wave. Every
// lane of the wave must take part.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
    int v = (int)x;
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)v;
}

// ---- launch geometry -----------------------------------------------------

constexpr int kBlock = 256;
constexpr int kXcds = 8;

// Maps the dispatch-order workgroup id onto a tile index so that each of the
// 8 XCDs (which take workgroups round-robin) owns one contiguous run of
// tiles: a run of cell-sorted particles shares its neighbours, and this keeps
// their gathers in the XCD's own L2. A bijection on [0, nblocks).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks) {
    if (nblocks < 2u * kXcds) return b;
    uint32_t per = (nblocks + kXcds - 1) / kXcds;
    uint32_t tall = nblocks % kXcds;
    if (tall == 0) tall = kXcds;
    uint32_t xcd = b % kXcds, local = b / kXcds;
    if (xcd < tall) return xcd * per + local;
    return tall * per + (xcd - tall) * (per - 1) + local;
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace fluid
